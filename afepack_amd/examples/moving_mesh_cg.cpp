// moving_mesh_cg.cpp -- a driver written against the drop-in AFEPack classes:
// a variable-coefficient elliptic problem on a mesh whose interior vertices
// move between solves, solved by a conjugate-gradient loop written here on
// host vectors, preconditioned by AMGPreconditioner::vmult.  The operator is
// a user class overriding getElementMatrix (evaluated on the host, gathered
// on the device); the preconditioner keeps its interpolation across the moves
// and only refreshes its Galerkin products (AMGPreconditioner::reinit =
// AMGSolver::lazyReinit on the same sparsity pattern).
//
// Problem: -div(a grad u) = f on the unit square, u = 0 on the boundary,
// a(x, y) = 1 + x y, exact u = sin(pi x) sin(pi y).
//
// usage: moving_mesh_cg <n> [moves]
// prints, per mesh position, the CG iterations and the L2 / H1 errors; the
// last line is "final L2 error = <value>".
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include <AFEPack/AMGSolver.h>
#include <AFEPack/BilinearOperator.h>
#include <AFEPack/FEMSpace.h>
#include <AFEPack/Functional.h>
#include <AFEPack/Operator.h>

static double coef(const double *p) { return 1.0 + p[0] * p[1]; }
static double exact_u(const double *p) { return std::sin(M_PI * p[0]) * std::sin(M_PI * p[1]); }
static std::vector<double> exact_grad(const double *p) {
  return {M_PI * std::cos(M_PI * p[0]) * std::sin(M_PI * p[1]), M_PI * std::sin(M_PI * p[0]) * std::cos(M_PI * p[1])};
}
// f = -(a u_xx + a u_yy + a_x u_x + a_y u_y) with a_x = y, a_y = x
static double source(const double *p) {
  const std::vector<double> g = exact_grad(p);
  return 2.0 * M_PI * M_PI * coef(p) * exact_u(p) - (p[1] * g[0] + p[0] * g[1]);
}
static double zero(const double *) { return 0.0; }

// stiffness with the coefficient a evaluated at the quadrature points
struct CoefStiff : public BilinearOperator<2, double> {
  explicit CoefStiff(FEMSpace<double, 2> &sp) : BilinearOperator<2, double>(sp, sp) {}
  void getElementMatrix(const Element<double, 2> &e0, const Element<double, 2> &e1,
                        const ActiveElementPairIterator<2>::State = ActiveElementPairIterator<2>::EQUAL) override {
    const int nd = (int)e0.dof().size();
    (void)e1;
    const double volume = e0.templateElement().volume();
    const auto qi = e0.findQuadratureInfo(algebricAccuracy());
    const auto jac = e0.local_to_global_jacobian(qi.quadraturePoint());
    const auto q = e0.local_to_global(qi.quadraturePoint());
    const auto grad = e0.basis_function_gradient(q);
    for (int l = 0; l < qi.n_quadraturePoint(); l++) {
      const double w = qi.weight(l) * jac[l] * volume * coef(q[l]);
      for (int j = 0; j < nd; j++)
        for (int k = 0; k < nd; k++) elementMatrix(j, k) += w * innerProduct(grad[j][l], grad[k][l]);
    }
  }
};

static double dot(const Vector<double> &a, const Vector<double> &b) {
  double s = 0.0;
  for (size_t i = 0; i < a.size(); i++) s += a(i) * b(i);
  return s;
}

// preconditioned CG on host vectors; returns the iterations used.  The
// preconditioner (forward Gauss-Seidel on both sides of the V-cycle) is not
// symmetric, so beta is the flexible (Polak-Ribiere) form z1.(r1 - r0) / (r0.z0).
static int pcg(const SparseMatrix<double> &A, const AMGPreconditioner &M, Vector<double> &x, const Vector<double> &b,
               double rtol, int max_it) {
  const size_t n = b.size();
  Vector<double> r(n), z(n), p(n), q(n), r0(n);
  A.vmult(q, x);
  for (size_t i = 0; i < n; i++) r(i) = b(i) - q(i);
  const double bnorm = std::sqrt(dot(b, b));
  z = 0.0;
  M.vmult(z, r);
  p = z;
  double rz = dot(r, z);
  for (int it = 1; it <= max_it; it++) {
    A.vmult(q, p);
    const double alpha = rz / dot(p, q);
    r0 = r;
    for (size_t i = 0; i < n; i++) {
      x(i) += alpha * p(i);
      r(i) -= alpha * q(i);
    }
    if (std::sqrt(dot(r, r)) <= rtol * bnorm) return it;
    z = 0.0;
    M.vmult(z, r);
    double zdr = 0.0;
    for (size_t i = 0; i < n; i++) zdr += z(i) * (r(i) - r0(i));
    for (size_t i = 0; i < n; i++) p(i) = z(i) + (zdr / rz) * p(i);
    rz = dot(r, z);
  }
  return -1;
}

int main(int argc, char **argv) {
  const int n = argc > 1 ? atoi(argv[1]) : 32;
  const int moves = argc > 2 ? atoi(argv[2]) : 3;
  if (n < 2 || moves < 1) {
    std::fprintf(stderr, "usage: moving_mesh_cg <n >= 2> [moves >= 1]\n");
    return 1;
  }
  try {
    Mesh<2, 2> mesh;
    buildSquareMesh(mesh, n, 0.15, 2027);
    const std::vector<double> base = mesh.coords;

    TemplateGeometry<2> geo;
    CoordTransform<2, 2> crd;
    TemplateDOF<2> tdof(geo);
    BasisFunctionAdmin<double, 2, 2> bas(tdof);
    geo.readData("triangle.tmp_geo");
    crd.readData("triangle.crd_trs");
    tdof.readData("triangle.2.tmp_dof");
    bas.readData("triangle.2.bas_fun");
    std::vector<TemplateElement<double, 2, 2>> tmpl(1);
    tmpl[0].reinit(geo, tdof, crd, bas);

    FEMSpace<double, 2> sp(mesh, tmpl);
    sp.element().resize(mesh.n_geometry(2));
    for (int i = 0; i < mesh.n_geometry(2); i++) sp.element(i).reinit(sp, i, 0);
    sp.buildElement();
    sp.buildDof();
    sp.buildDofBoundaryMark();

    const int acc = 6;
    CoefStiff A(sp);
    A.algebricAccuracy() = acc;
    AMGPreconditioner pre;
    FunctionFunction<double> u(&exact_u, &exact_grad);
    double l2 = 0.0;
    for (int m = 0; m < moves; m++) {
      // interior vertices displaced by 0.1 h sin(2 pi k x) from the base mesh
      mesh.coords = base;
      if (m > 0)
        detail::check(afe_mesh_displace(2, mesh.n_point(), mesh.coords.data(), mesh.vmark.data(), 0.1 / n, m));
      A.build();
      Vector<double> f;
      Operator::L2Discretize(&source, sp, f, acc);
      FEMFunction<double, 2> x(sp);
      BoundaryFunction<double, 2> bc(BoundaryConditionInfo::DIRICHLET, 1, &zero);
      BoundaryConditionAdmin<double, 2> admin(sp);
      admin.add(bc);
      admin.apply(A, x, f);
      pre.reinit(A, 1); // lazyInit on the first mesh, lazyReinit afterwards
      const int its = pcg(A, pre, x, f, 1.0e-10, 500);
      l2 = Functional::L2Error(x, u, acc);
      const double h1 = Functional::H1SemiError(x, u, acc);
      std::printf("mesh %d: %d DOFs, pcg %d iterations, L2 %.6e, H1 %.6e\n", m, sp.n_dof(), its, l2, h1);
      if (its < 0) {
        std::fprintf(stderr, "CG did not converge on mesh %d\n", m);
        return 4;
      }
    }
    std::printf("final L2 error = %.6e\n", l2);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "moving_mesh_cg failed: %s\n", e.what());
    return 3;
  }
  return 0;
}
