// facade_check.cpp -- runs the assemble-and-solve path through the drop-in
// C++ classes and writes every intermediate result as raw little-endian
// arrays, so that they can be compared with an independent computation.
//
// usage: facade_check <kind 1|2|3> <n> <outdir> [easymesh-base]
//   kind 1 = triangle.1 (P1), 2 = triangle.2 (P2), 3 = tetrahedron.1; with an
//   EasyMesh base name (2D) the mesh is read from <base>.n/.s/.e instead of
//   generated
// files: rp.i64 col.i64 (pattern), K.f64 (StiffMatrix), KH.f64 (the same
// form evaluated on the host by a derived class), K2.f64 (an override that
// doubles it), M.f64 (MassMatrix), f.f64 (L2Discretize), Ka.f64 fa.f64 ua.f64
// (after BoundaryConditionAdmin::apply), g.f64 (Dirichlet data per DOF),
// x.f64 (AMGSolver::solve), xp.f64 (AMG-preconditioned CG), xv.f64
// (AMGPreconditioner::vmult of f from zero), err.txt (L2Error, H1SemiError of
// x against prod sin(pi x_d)).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>

#include <AFEPack/AMGSolver.h>
#include <AFEPack/BilinearOperator.h>
#include <AFEPack/FEMSpace.h>
#include <AFEPack/Functional.h>
#include <AFEPack/Operator.h>

static int g_dim = 2;

double sinprod(const double *p) {
  double v = 1.0;
  for (int d = 0; d < g_dim; d++) v *= sin(M_PI * p[d]);
  return v;
}
std::vector<double> sinprod_grad(const double *p) {
  std::vector<double> g(g_dim);
  for (int d = 0; d < g_dim; d++) {
    double v = M_PI * cos(M_PI * p[d]);
    for (int e = 0; e < g_dim; e++)
      if (e != d) v *= sin(M_PI * p[e]);
    g[d] = v;
  }
  return g;
}
double poisson_f(const double *p) { return g_dim * M_PI * M_PI * sinprod(p); }
double dirichlet_g(const double *p) { return 1.0 + 0.5 * p[0] - 0.25 * p[1] + sin(3.0 * p[0]) * cos(2.0 * p[1]); }

template <class T> static void dump(const std::string &path, const T *p, size_t n) {
  std::ofstream os(path, std::ios::binary);
  os.write((const char *)p, n * sizeof(T));
  if (!os) {
    std::cerr << "cannot write " << path << std::endl;
    exit(2);
  }
}
static void dump_vec(const std::string &path, const Vector<double> &v) { dump(path, v.data(), v.size()); }
static void dump_mat(const std::string &path, const SparseMatrix<double> &a) {
  dump(path, a.values().data(), a.values().size());
}

template <int DIM> struct HostStiff : public StiffMatrix<DIM, double> {
  HostStiff(FEMSpace<double, DIM> &sp) : StiffMatrix<DIM, double>(sp) {}
};

template <int DIM> struct DoubledStiff : public StiffMatrix<DIM, double> {
  DoubledStiff(FEMSpace<double, DIM> &sp) : StiffMatrix<DIM, double>(sp) {}
  void getElementMatrix(const Element<double, DIM> &e0, const Element<double, DIM> &e1,
                        const typename ActiveElementPairIterator<DIM>::State s =
                            ActiveElementPairIterator<DIM>::EQUAL) override {
    StiffMatrix<DIM, double>::getElementMatrix(e0, e1, s);
    DenseMatrix<double> &m = this->elementMatrix();
    for (size_t i = 0; i < m.m(); i++)
      for (size_t j = 0; j < m.n(); j++) m(i, j) *= 2.0;
  }
};

template <int DIM> static int run(int kind, int n, const std::string &out, const std::string &easymesh) {
  g_dim = DIM;
  Mesh<DIM, DIM> mesh;
  if constexpr (DIM == 2) {
    if (!easymesh.empty()) {
      EasyMesh em;
      em.readData(easymesh);
      mesh.coords = em.coords;
      mesh.elem = em.elem;
      mesh.vmark = em.vmark;
    } else {
      buildSquareMesh(mesh, n, 0.2, 7);
    }
  } else {
    buildCubeMesh(mesh, n, 0.05, 7);
  }

  TemplateGeometry<DIM> geo;
  CoordTransform<DIM, DIM> crd;
  TemplateDOF<DIM> tdof(geo);
  BasisFunctionAdmin<double, DIM, DIM> bas(tdof);
  const std::string base = DIM == 2 ? "triangle" : "tetrahedron";
  const std::string order = kind == AFE_P2_TRI ? ".2" : ".1";
  geo.readData(base + ".tmp_geo");
  crd.readData(base + ".crd_trs");
  tdof.readData(base + order + ".tmp_dof");
  bas.readData(base + order + ".bas_fun");
  std::vector<TemplateElement<double, DIM, DIM>> tmpl(1);
  tmpl[0].reinit(geo, tdof, crd, bas);

  FEMSpace<double, DIM> sp(mesh, tmpl);
  const int n_element = mesh.n_geometry(DIM);
  sp.element().resize(n_element);
  for (int i = 0; i < n_element; i++) sp.element(i).reinit(sp, i, 0);
  sp.buildElement();
  sp.buildDof();
  sp.buildDofBoundaryMark();
  const int acc = 4;

  StiffMatrix<DIM, double> K(sp);
  K.algebricAccuracy() = acc;
  K.build();
  const SparsityPattern &pat = K.getSparsityPattern();
  std::vector<int64_t> rp(pat.row_ptr().begin(), pat.row_ptr().end()), col(pat.col_indices().begin(), pat.col_indices().end());
  dump(out + "/rp.i64", rp.data(), rp.size());
  dump(out + "/col.i64", col.data(), col.size());
  dump_mat(out + "/K.f64", K);

  HostStiff<DIM> KH(sp);
  KH.algebricAccuracy() = acc;
  KH.build();
  dump_mat(out + "/KH.f64", KH);

  DoubledStiff<DIM> K2(sp);
  K2.algebricAccuracy() = acc;
  K2.build();
  dump_mat(out + "/K2.f64", K2);

  MassMatrix<DIM, double> M(sp);
  M.algebricAccuracy() = acc;
  M.build();
  dump_mat(out + "/M.f64", M);

  Vector<double> f;
  Operator::L2Discretize(&poisson_f, sp, f, acc);
  dump_vec(out + "/f.f64", f);

  std::vector<double> g(sp.n_dof());
  for (int i = 0; i < sp.n_dof(); i++) g[i] = dirichlet_g(sp.dofInterpPoint(i));
  dump(out + "/g.f64", g.data(), g.size());
  BoundaryFunction<double, DIM> bc(BoundaryConditionInfo::DIRICHLET, 1, &dirichlet_g);
  BoundaryConditionAdmin<double, DIM> admin(sp);
  admin.add(bc);
  FEMFunction<double, DIM> x(sp);
  admin.apply(K, x, f);
  dump_mat(out + "/Ka.f64", K);
  dump_vec(out + "/fa.f64", f);
  dump_vec(out + "/ua.f64", x);

  AMGSolver solver(K);
  solver.solve(x, f, 1.0e-12, 60);
  dump_vec(out + "/x.f64", x);

  FEMFunction<double, DIM> xp(sp);
  const int its = solver.pcg(xp, f, 1.0e-11, 300, true);
  dump_vec(out + "/xp.f64", xp);

  AMGPreconditioner pre(K, 2, 3);
  Vector<double> xv(sp.n_dof());
  pre.vmult(xv, f);
  dump_vec(out + "/xv.f64", xv);

  FunctionFunction<double> uex(&sinprod, &sinprod_grad);
  const double l2 = Functional::L2Error(xp, uex, acc), h1 = Functional::H1SemiError(xp, uex, acc);
  std::ofstream es(out + "/err.txt");
  es.precision(17);
  es << l2 << " " << h1 << " " << its << "\n";
  std::printf("facade_check kind %d n %d: %d DOFs, pcg %d its, L2 %.6e H1 %.6e\n", kind, n, sp.n_dof(), its, l2, h1);
  return 0;
}

int main(int argc, char **argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: facade_check <kind 1|2|3> <n> <outdir>\n");
    return 1;
  }
  const int kind = atoi(argv[1]), n = atoi(argv[2]);
  try {
    const std::string em = argc > 4 ? argv[4] : "";
    return kind == AFE_P1_TET ? run<3>(kind, n, argv[3], em) : run<2>(kind, n, argv[3], em);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "facade_check failed: %s\n", e.what());
    return 3;
  }
}
