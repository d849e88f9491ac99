// AMGSolver.h -- drop-in AMGSolver / AMGPreconditioner (include/AFEPack/
// AMGSolver.h:66-196, src/AMGSolver.cpp:384-699 of the reference).
//
// The hierarchy is built by the library from the matrix's values (the
// reference's coarsening, on the host, uploaded once); solve() and vmult()
// run the V-cycles on the device.  Host vectors are copied to HBM and back
// around each call; callers holding device vectors use the C ABI directly
// (afe_amg_solve / afe_amg_pcg).
#pragma once
#include "FEMSpace.h"

namespace AFEPack {

class AMGSolver {
public:
  typedef SparseMatrix<double> Matrix;
  typedef unsigned int u_int;

  AMGSolver() = default;
  AMGSolver(const Matrix &M, double tol = 1.0e-12, u_int s = 3, u_int nmpo = 50, double nmpsd = 0.382,
            double alp = 0.25)
      : toler_(tol), smooth_(s), nmpo_(nmpo), nmpsd_(nmpsd), alpha_(alp) {
    reinit(M);
  }
  virtual ~AMGSolver() { clear(); }
  AMGSolver(const AMGSolver &) = delete;
  AMGSolver &operator=(const AMGSolver &) = delete;

  void clear() {
    if (g_) afe_amg_destroy(g_);
    if (own_) afe_matrix_destroy(own_);
    g_ = nullptr;
    own_ = nullptr;
    pat_ = nullptr;
  }
  double tolerence() const { return toler_; }
  double &tolerence() { return toler_; }
  u_int smoothStep() const { return smooth_; }
  u_int &smoothStep() { return smooth_; }
  bool &isSolveMostProjectExactly() { return exact_coarse_; }
  bool isSolveMostProjectExactly() const { return exact_coarse_; }

  // AMGSolver::reinit: full coarsening (src/AMGSolver.cpp:384-438)
  void reinit(const Matrix &M) { build(M, 0); }
  // AMGSolver::lazyReinit (src/AMGSolver.cpp:505-545).  The first call builds
  // the lazy hierarchy (lazyInit).  Later calls must pass a matrix on the
  // same SparsityPattern object: P is kept and the Galerkin products are
  // recomputed from the new values; another pattern prints the reference's
  // message and aborts.  (After reinit, whose hierarchy has no kept P^T,
  // the reference's refresh reads past its arrays; here it builds the lazy
  // hierarchy anew.)
  void lazyReinit(const Matrix &M) {
    if (!g_) {
      build(M, 1);
      return;
    }
    if (&M.get_sparsity_pattern() != pat_) detail::fatal("The solver is not initialized with the same sparsity pattern.");
    if (!lazy_) {
      build(M, 1);
      return;
    }
    M.push_to_device();
    afe_matrix *a = M.device_matrix();
    if (!a && own_) {
      // host-assembled values: refresh the device copy the hierarchy was
      // built on in place, so the hierarchy never points at a freed matrix
      detail::check(afe_matrix_set_values(own_, M.values().data()));
      a = own_;
    }
    const int rc = a ? afe_amg_lazy_reinit(g_, a) : AFE_E_PATTERN;
    if (rc == AFE_E_PATTERN) {
      // same pattern object, but the matrix is now bound to another device
      // copy: lazyProject depends on the pattern only, so a fresh lazy build
      // gives the same P and the same Galerkin products
      build(M, 1);
      return;
    }
    detail::check(rc);
  }
  // AMGSolver::solve (src/AMGSolver.cpp:549-639): exactly step+1 V(s,s)
  // cycles, mode 0 reports the final residual on stderr
  void solve(Vector<double> &x, const Vector<double> &r, double tol = 0.0, size_t step = 20, int mode = 0) const {
    prepare(x, r);
    detail::DeviceVector dx(x.size()), dr(r.size());
    dx.upload(x.data());
    dr.upload(r.data());
    afe_solve_stats st;
    detail::check(afe_amg_solve(g_, dx.p, dr.p, tol == 0.0 ? toler_ : tol, (int64_t)step, mode, 0.0, &st));
    dx.download(x.data());
  }
  // FP64 conjugate gradients preconditioned by one symmetric V-cycle (FP32
  // or FP64) per iteration; returns the iteration count (no counterpart in
  // the reference, which pairs AMGPreconditioner with an external CG)
  int pcg(Vector<double> &x, const Vector<double> &b, double rtol, int max_it, bool fp32 = true) const {
    prepare(x, b);
    detail::DeviceVector dx(x.size()), db(b.size());
    dx.upload(x.data());
    db.upload(b.data());
    afe_solve_stats st;
    detail::check(afe_amg_pcg(g_, dx.p, db.p, rtol, max_it, fp32 ? AFE_PREC_FP32 : AFE_PREC_FP64, &st));
    dx.download(x.data());
    return st.iterations;
  }
  afe_amg *handle() const { return g_; }

protected:
  void prepare(const Vector<double> &x, const Vector<double> &r) const {
    if (!g_) throw std::runtime_error("AMGSolver used before reinit");
    if (x.size() != r.size() || x.size() != n_) throw std::runtime_error("AMGSolver: size mismatch");
    if (smooth_ != smooth_set_) {
      detail::check(afe_amg_set_smooth(g_, (int32_t)smooth_));
      smooth_set_ = smooth_;
    }
    if (exact_coarse_ != exact_set_) {
      detail::check(afe_amg_set_coarse_exact(g_, exact_coarse_ ? 1 : 0));
      exact_set_ = exact_coarse_;
    }
  }
  void build(const Matrix &M, int lazy) {
    clear();
    M.push_to_device();
    afe_matrix *a = M.device_matrix();
    if (!a) {
      // a host-assembled matrix: a device copy of its CSR, kept for the
      // lifetime of the hierarchy (the hierarchy borrows its fine matrix)
      const auto &pat = M.get_sparsity_pattern();
      std::vector<int64_t> rp(pat.row_ptr().begin(), pat.row_ptr().end()),
          col(pat.col_indices().begin(), pat.col_indices().end());
      detail::check(
          afe_matrix_create_csr(detail::device(), (int64_t)M.m(), rp.data(), col.data(), M.values().data(), &own_));
      a = own_;
    }
    afe_amg_params p;
    p.alpha = alpha_;
    p.nmpo = nmpo_;
    p.nmpsd = nmpsd_;
    p.lazy = lazy;
    p.smooth = (int32_t)smooth_;
    p.build_fp32 = 1;
    detail::check(afe_amg_create(a, &p, &g_));
    lazy_ = lazy;
    pat_ = &M.get_sparsity_pattern();
    n_ = M.m();
    smooth_set_ = smooth_;
    exact_set_ = true;
  }

  afe_amg *g_ = nullptr;
  afe_matrix *own_ = nullptr;
  const SparsityPattern *pat_ = nullptr; // the pattern object of the fine matrix
  size_t n_ = 0;
  int lazy_ = 0;
  double toler_ = 1.0e-12;
  u_int smooth_ = 3, nmpo_ = 50;
  mutable u_int smooth_set_ = 3;
  double nmpsd_ = 0.382, alpha_ = 0.25;
  bool exact_coarse_ = true;
  mutable bool exact_set_ = true;
};

// AMGSolver used as a preconditioner: vmult = iterate_step cycles from the
// given x in precondition mode (AMGSolver.h:171-196)
class AMGPreconditioner : public AMGSolver {
public:
  AMGPreconditioner() = default;
  AMGPreconditioner(const Matrix &M, u_int is = 5, u_int s = 3) : iterate_step_(is) {
    smoothStep() = s;
    lazyReinit(M);
  }
  void reinit(const Matrix &M, u_int is = 5) {
    AMGSolver::lazyReinit(M);
    iterate_step_ = is;
  }
  u_int &iterateStep() { return iterate_step_; }
  const u_int &iterateStep() const { return iterate_step_; }
  void vmult(Vector<double> &x, const Vector<double> &b) const {
    prepare(x, b);
    detail::DeviceVector dx(x.size()), db(b.size());
    dx.upload(x.data());
    db.upload(b.data());
    detail::check(afe_amg_precondition(g_, dx.p, db.p, (int32_t)iterate_step_));
    dx.download(x.data());
  }

private:
  u_int iterate_step_ = 5;
};

} // namespace AFEPack

using namespace AFEPack;
