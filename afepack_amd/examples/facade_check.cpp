// facade_check.cpp -- runs the assemble-and-solve path through the drop-in
// C++ classes and writes every intermediate result as raw little-endian
// arrays, so that they can be compared with an independent computation.
//
// usage: facade_check <kind 1|2|3> <n> <outdir> [easymesh-base]
//   kind 1 = triangle.1 (P1), 2 = triangle.2 (P2), 3 = tetrahedron.1; with an
//   EasyMesh base name (2D) the mesh is read from <base>.n/.s/.e instead of
//   generated, with a name ending in .mesh from that file in the Mesh text
//   format
// files: rp.i64 col.i64 (pattern), K.f64 (StiffMatrix), KH.f64 (the same
// form evaluated on the host by a derived class), K2.f64 (an override that
// doubles it), M.f64 (MassMatrix), f.f64 (L2Discretize), Ka.f64 fa.f64 ua.f64
// (after BoundaryConditionAdmin::apply), Kn.f64 fn.f64 (after apply with
// preserve_symmetry = false), g.f64 (Dirichlet data per DOF),
// x.f64 (AMGSolver::solve), xp.f64 (AMG-preconditioned CG), xv.f64
// (AMGPreconditioner::vmult of f from zero), err.txt (L2Error, H1SemiError of
// x against prod sin(pi x_d)), L2.f64 (L2InnerProduct on one space), KM.f64
// (a user operator declared BilinearOperator<DIM, double>: stiffness plus
// mass, evaluated on the host), info.txt (nMaxCouplingDof), xg.f64
// (AMGSolver::solve with isSolveMostProjectExactly() = false), cm.f64 and
// Km.f64 (moved vertex coordinates and the stiffness re-assembled on them),
// lz1.f64 (level-1 values of a lazy hierarchy after lazyReinit on the moved
// matrix), hl1.f64 (level-1 values after two lazyReinit calls on a
// host-assembled matrix with values 3 K), foreign.txt (1 when lazyReinit with
// another pattern object reached the reference's abort), L2x_rp.i64
// L2x_col.i64 L2x.f64 (L2InnerProduct of this space and a second one on the
// same mesh: the other order in 2D, another P1 space in 3D), L2xd.f64 (a
// derived two-space operator doubling it, evaluated on the host), infox.txt
// (its nMaxCouplingDof, nDof0, nDof1, m, n).
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

// a user operator written against the reference's default second value
// type (BilinearOperator<DIM, double>): stiffness plus mass in one loop
template <int DIM> struct StiffPlusMass : public BilinearOperator<DIM, double> {
  StiffPlusMass(FEMSpace<double, DIM> &sp) : BilinearOperator<DIM, double>(sp, sp) {}
  void getElementMatrix(const Element<double, DIM> &e0, const Element<double, DIM> &e1,
                        const typename ActiveElementPairIterator<DIM>::State =
                            ActiveElementPairIterator<DIM>::EQUAL) override {
    const int n0 = (int)e0.dof().size(), n1 = (int)e1.dof().size();
    const double volume = e0.templateElement().volume();
    const auto qi = e0.findQuadratureInfo(this->algebricAccuracy());
    const auto jac = e0.local_to_global_jacobian(qi.quadraturePoint());
    const auto q = e0.local_to_global(qi.quadraturePoint());
    const auto bv = e0.basis_function_value(q);
    const auto bg = e0.basis_function_gradient(q);
    for (int l = 0; l < qi.n_quadraturePoint(); l++) {
      const double Jxw = qi.weight(l) * jac[l] * volume;
      for (int j = 0; j < n0; j++)
        for (int k = 0; k < n1; k++)
          this->elementMatrix(j, k) += Jxw * (innerProduct(bg[j][l], bg[k][l]) + bv[j][l] * bv[k][l]);
    }
  }
};

// a user operator on two spaces: L2InnerProduct's element matrix, doubled
template <int DIM> struct DoubledL2 : public L2InnerProduct<DIM, double> {
  DoubledL2(FEMSpace<double, DIM> &s0, FEMSpace<double, DIM> &s1) : L2InnerProduct<DIM, double>(s0, s1) {}
  void getElementMatrix(const Element<double, DIM> &e0, const Element<double, DIM> &e1,
                        const typename ActiveElementPairIterator<DIM>::State s =
                            ActiveElementPairIterator<DIM>::EQUAL) override {
    L2InnerProduct<DIM, double>::getElementMatrix(e0, e1, s);
    DenseMatrix<double> &m = this->elementMatrix();
    for (size_t i = 0; i < m.m(); i++)
      for (size_t j = 0; j < m.n(); j++) m(i, j) *= 2.0;
  }
};

struct ForeignPattern {};
static void throw_foreign(const char *) { throw ForeignPattern{}; }

// level-1 values of a hierarchy (afe_amg_level_csr)
static void dump_level1(const std::string &path, afe_amg *g) {
  int64_t m, n, nnz;
  detail::check(afe_amg_level_csr(g, 1, 0, &m, &n, &nnz, nullptr, nullptr, nullptr));
  std::vector<double> v((size_t)nnz);
  detail::check(afe_amg_level_csr(g, 1, 0, &m, &n, &nnz, nullptr, nullptr, v.data()));
  dump(path, v.data(), v.size());
}

template <int DIM> static int run(int kind, int n, const std::string &out, const std::string &easymesh) {
  g_dim = DIM;
  Mesh<DIM, DIM> mesh;
  if constexpr (DIM == 2) {
    if (easymesh.size() > 5 && easymesh.substr(easymesh.size() - 5) == ".mesh") {
      mesh.readData(easymesh); // the Mesh text format
    } else if (!easymesh.empty()) {
      EasyMesh em;
      em.readData(easymesh);
      mesh.coords = em.coords;
      mesh.elem = em.elem;
      mesh.vmark = em.vmark;
      mesh.side_vert = em.side_vert;
      mesh.side_mark = em.side_mark;
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
  {
    // preserve_symmetry = false on a copy of the system: the boundary rows only
    HostStiff<DIM> Kn(sp);
    Kn.algebricAccuracy() = acc;
    Kn.build();
    Vector<double> fn(f.size());
    for (size_t i = 0; i < f.size(); i++) fn(i) = f(i);
    FEMFunction<double, DIM> xn(sp);
    admin.apply(Kn, xn, fn, false);
    dump_mat(out + "/Kn.f64", Kn);
    dump_vec(out + "/fn.f64", fn);
  }
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

  // L2InnerProduct on one space: the device mass kernel
  L2InnerProduct<DIM, double> L2(sp, sp);
  L2.algebricAccuracy() = acc;
  L2.build();
  dump_mat(out + "/L2.f64", L2);
  StiffPlusMass<DIM> KM(sp);
  KM.algebricAccuracy() = acc;
  KM.build();
  dump_mat(out + "/KM.f64", KM);
  std::ofstream(out + "/info.txt") << KM.nMaxCouplingDof() << " " << KM.nDof0() << " " << KM.nDof1() << "\n";

  // two FE spaces on one mesh: the other order in 2D, a second P1 space of
  // tetrahedra in 3D; L2InnerProduct(sp, sp2) on the device, a derived
  // class that doubles it on the host
  TemplateDOF<DIM> tdof2(geo);
  BasisFunctionAdmin<double, DIM, DIM> bas2(tdof2);
  const std::string order2 = DIM == 3 ? ".1" : (kind == AFE_P2_TRI ? ".1" : ".2");
  tdof2.readData(base + order2 + ".tmp_dof");
  bas2.readData(base + order2 + ".bas_fun");
  std::vector<TemplateElement<double, DIM, DIM>> tmpl2(1);
  tmpl2[0].reinit(geo, tdof2, crd, bas2);
  FEMSpace<double, DIM> sp2(mesh, tmpl2);
  sp2.element().resize(n_element);
  for (int i = 0; i < n_element; i++) sp2.element(i).reinit(sp2, i, 0);
  sp2.buildElement();
  sp2.buildDof();
  sp2.buildDofBoundaryMark();
  L2InnerProduct<DIM, double> L2x(sp, sp2);
  L2x.algebricAccuracy() = acc;
  L2x.build();
  const SparsityPattern &px = L2x.getSparsityPattern();
  std::vector<int64_t> rpx(px.row_ptr().begin(), px.row_ptr().end()), colx(px.col_indices().begin(), px.col_indices().end());
  dump(out + "/L2x_rp.i64", rpx.data(), rpx.size());
  dump(out + "/L2x_col.i64", colx.data(), colx.size());
  dump_mat(out + "/L2x.f64", L2x);
  DoubledL2<DIM> L2xd(sp, sp2);
  L2xd.algebricAccuracy() = acc;
  L2xd.build();
  dump_mat(out + "/L2xd.f64", L2xd);
  std::ofstream(out + "/infox.txt") << L2x.nMaxCouplingDof() << " " << L2x.nDof0() << " " << L2x.nDof1() << " "
                                    << L2x.m() << " " << L2x.n() << "\n";

  // the coarsest level smoothed instead of solved exactly
  AMGSolver gsolver(K);
  gsolver.isSolveMostProjectExactly() = false;
  Vector<double> xg(sp.n_dof());
  gsolver.solve(xg, f, 1.0e-12, 400, 1);
  dump_vec(out + "/xg.f64", xg);

  if (n <= 0) { // a mesh read from a file: the generated-mesh checks end here
    std::printf("facade_check kind %d file mesh: %d DOFs, pcg %d its, L2 %.6e\n", kind, sp.n_dof(), its, l2);
    return 0;
  }
  // moving mesh: a lazy hierarchy on K, the interior vertices displaced,
  // K re-assembled on its own pattern object, lazyReinit
  AMGSolver lz;
  lz.lazyReinit(K);
  {
    double h = 1.0 / n;
    detail::check(afe_mesh_displace(DIM, mesh.n_point(), mesh.coords.data(), mesh.vmark.data(), 0.1 * h, 2));
  }
  dump(out + "/cm.f64", mesh.coords.data(), mesh.coords.size());
  K.build();
  dump_mat(out + "/Km.f64", K);
  lz.lazyReinit(K);
  dump_level1(out + "/lz1.f64", lz.handle());
  // another pattern object: the reference prints its message and aborts
  int foreign = 0;
  detail::fatal_hook() = &throw_foreign;
  try {
    lz.lazyReinit(M);
  } catch (const ForeignPattern &) {
    foreign = 1;
  }
  detail::fatal_hook() = nullptr;
  std::ofstream(out + "/foreign.txt") << foreign << "\n";

  // a host-assembled matrix (no device binding) refreshed twice in place
  SparsityPattern hpat;
  {
    const SparsityPattern &kp = K.getSparsityPattern();
    hpat.assign_compressed(kp.n_rows(), kp.n_cols(), kp.row_ptr(), kp.col_indices());
  }
  SparseMatrix<double> HM(hpat);
  const std::vector<double> kv = K.values();
  for (size_t k = 0; k < kv.size(); k++) HM.global_entry(k) = kv[k];
  AMGSolver hl;
  hl.lazyReinit(HM);
  for (int pass = 2; pass <= 3; pass++) {
    for (size_t k = 0; k < kv.size(); k++) HM.global_entry(k) = pass * kv[k];
    hl.lazyReinit(HM);
  }
  dump_level1(out + "/hl1.f64", hl.handle());

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
