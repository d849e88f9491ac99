// BilinearOperator.h -- drop-in BilinearOperator / StiffMatrix / MassMatrix
// (include/AFEPack/BilinearOperator.h:76-224 and
// BilinearOperator.templates.h:55-255,300-406 of the reference).
//
// build() = buildSparsityPattern() + buildSparseMatrix().  On one space the
// pattern is the one afe_space_create computed (first-touch numbering,
// diagonal first, rest ascending); on two spaces of one mesh it is the
// rectangular pattern of afe_matrix_create_pair (the reference's same-mesh
// branch; spaces on different meshes are not supported).
// buildSparseMatrix() assembles on the device:
//   * StiffMatrix and MassMatrix objects run the device element kernels
//     (afe_assemble, op STIFF / MASS);
//   * a derived class -- the reference's extension point, which overrides
//     getElementMatrix (e.g. a variable-coefficient stiffness) -- is
//     evaluated on the host element by element, as the reference does, and
//     the element matrices are gathered into the CSR values on the device
//     (afe_assemble_elements).
// The matrix keeps a device copy; the host values are copied back lazily on
// the first host access.
#pragma once
#include <algorithm>
#include <typeinfo>

#include "FEMSpace.h"

namespace AFEPack {

template <int DIM, class value_type0, class value_type1 = value_type0, int DOW = DIM, int TDIM0 = DIM,
          int TDIM1 = DIM>
class BilinearOperator : public SparseMatrix<double> {
public:
  using space0_t = FEMSpace<value_type0, DIM, DOW, TDIM0>;
  using space1_t = FEMSpace<value_type1, DIM, DOW, TDIM1>;
  using element0_t = Element<value_type0, DIM, DOW, TDIM0>;
  using element1_t = Element<value_type1, DIM, DOW, TDIM1>;

  BilinearOperator() = default;
  BilinearOperator(space0_t &s0, space1_t &s1) { reinit(s0, s1); }
  virtual ~BilinearOperator() {
    if (pair_dev_) afe_matrix_destroy(pair_dev_);
  }
  void reinit(space0_t &s0, space1_t &s1) {
    sp0_ = &s0;
    sp1_ = &s1;
  }
  const space0_t &FEMSpace0() const { return *sp0_; }
  const space1_t &FEMSpace1() const { return *sp1_; }
  space0_t *FEMSpace0() { return sp0_; }
  space1_t *FEMSpace1() { return sp1_; }
  const SparsityPattern &getSparsityPattern() const { return pattern_store_; }
  SparsityPattern &getSparsityPattern() { return pattern_store_; }

  // BilinearOperator.h:101-116: sizes, the coupling bound of the pattern and
  // the DOFs of the element pair being assembled
  const int &nDof0() const { return n_dof0_; }
  const int &nDof1() const { return n_dof1_; }
  const int &nMaxCouplingDof() const { return n_max_coupling_; }
  int &nDof0() { return n_dof0_; }
  int &nDof1() { return n_dof1_; }
  int &nMaxCouplingDof() { return n_max_coupling_; }
  const std::vector<int> &elementDof0() const { return *element_dof0_; }
  const std::vector<int> &elementDof1() const { return *element_dof1_; }
  const int &elementDof0(const int &i) const { return (*element_dof0_)[i]; }
  const int &elementDof1(const int &i) const { return (*element_dof1_)[i]; }
  void getElementPattern(const element0_t &e0, const element1_t &e1) {
    element_dof0_ = &e0.dof();
    element_dof1_ = &e1.dof();
  }

  const DenseMatrix<double> &elementMatrix() const { return element_matrix_; }
  DenseMatrix<double> &elementMatrix() { return element_matrix_; }
  double elementMatrix(int i, int j) const { return element_matrix_(i, j); }
  double &elementMatrix(int i, int j) { return element_matrix_(i, j); }
  const int &algebricAccuracy() const { return acc_; }
  int &algebricAccuracy() { return acc_; }

  virtual void getElementMatrix(const element0_t &, const element1_t &,
                                const typename ActiveElementPairIterator<DIM>::State state =
                                    ActiveElementPairIterator<DIM>::EQUAL) = 0;
  virtual void build() {
    buildSparsityPattern();
    buildSparseMatrix();
  }
  // the CSR pattern (BilinearOperator.templates.h:64-172)
  virtual void buildSparsityPattern() {
    afe_space *h = sp0_->handle();
    if (!h) throw std::runtime_error("buildDof() must run before build()");
    if (!same_space()) {
      build_pair_pattern();
      return;
    }
    int64_t nd, nnz;
    int32_t nl;
    check(afe_space_info(h, &nd, &nnz, &nl));
    std::vector<int64_t> rp(nd + 1), col(nnz);
    check(afe_space_pattern(h, rp.data(), col.data()));
    pattern_store_.assign_compressed((size_t)nd, (size_t)nd, std::vector<size_t>(rp.begin(), rp.end()),
                                     std::vector<size_t>(col.begin(), col.end()));
    // buildDofInfo (BilinearOperator.templates.h:115-172): the largest sum of
    // element DOF counts over the elements sharing a DOF, capped at nDof1
    n_dof0_ = n_dof1_ = (int)nd;
    std::vector<int> coupling((size_t)nd, 0);
    for (int e = 0; e < sp0_->n_element(); e++)
      for (int d : sp0_->element(e).dof()) coupling[d] += (int)sp0_->element(e).dof().size();
    n_max_coupling_ = nd ? std::min(*std::max_element(coupling.begin(), coupling.end()), (int)nd) : 0;
  }
  // values on the device (BilinearOperator.templates.h:176-255)
  virtual void buildSparseMatrix() {
    if (!same_space()) {
      build_pair_matrix();
      return;
    }
    SparseMatrix<double>::reinit(pattern_store_);
    afe_matrix *a = nullptr;
    check(afe_matrix_create(sp0_->handle(), &a));
    bind_device(a, true);
    const int n_elem = sp0_->n_element();
    // the device element kernels serve exactly StiffMatrix / MassMatrix;
    // a derived class (which may override getElementMatrix) is evaluated on
    // the host element by element
    const int op = device_operator();
    const bool builtin = op >= 0;
    if (builtin) {
      check(afe_assemble(a, op, acc_, nullptr));
    } else {
      const int nl = n_elem ? (int)sp0_->element(0).dof().size() : 0;
      std::vector<double> em((size_t)n_elem * nl * nl);
      for (int e = 0; e < n_elem; e++) {
        getElementPattern(sp0_->element(e), sp1_->element(e));
        element_matrix_.reinit(nl, nl);
        getElementMatrix(sp0_->element(e), sp1_->element(e));
        std::copy(element_matrix_.data(), element_matrix_.data() + (size_t)nl * nl, em.begin() + (size_t)e * nl * nl);
      }
      check(afe_assemble_elements(a, em.data()));
    }
    mark_device_newer();
  }

protected:
  // AFE_OP_* of a built-in element matrix, -1 for user operators
  virtual int device_operator() const { return -1; }
  space0_t *sp0_ = nullptr;
  space1_t *sp1_ = nullptr;

private:
  bool same_space() const { return (const void *)sp0_ == (const void *)sp1_; }
  // two spaces: the device matrix owns the pattern; it is made here and the
  // values are filled by build_pair_matrix
  void build_pair_pattern() {
    if ((const void *)&sp0_->mesh() != (const void *)&sp1_->mesh())
      throw std::runtime_error("BilinearOperator: FE spaces on different meshes are not supported on this path");
    afe_space *h1 = sp1_->handle();
    if (!h1) throw std::runtime_error("buildDof() must run before build()");
    afe_matrix *a = nullptr;
    check(afe_matrix_create_pair(sp0_->handle(), h1, &a));
    int64_t m, n, nnz;
    int32_t mc;
    check(afe_matrix_info(a, &m, &n, &nnz, &mc));
    std::vector<int64_t> rp(m + 1), col(nnz);
    check(afe_matrix_pattern(a, rp.data(), col.data()));
    pattern_store_.assign_compressed((size_t)m, (size_t)n, std::vector<size_t>(rp.begin(), rp.end()),
                                     std::vector<size_t>(col.begin(), col.end()));
    n_dof0_ = (int)m;
    n_dof1_ = (int)n;
    n_max_coupling_ = mc;
    if (pair_dev_) afe_matrix_destroy(pair_dev_);
    pair_dev_ = a;
  }
  void build_pair_matrix() {
    if (!pair_dev_) build_pair_pattern();
    afe_matrix *a = pair_dev_;
    pair_dev_ = nullptr;
    SparseMatrix<double>::reinit(pattern_store_);
    bind_device(a, true);
    const int op = device_operator();
    if (op == AFE_OP_MASS) {
      check(afe_assemble(a, op, acc_, nullptr));
    } else {
      if (op >= 0) throw std::runtime_error("BilinearOperator: this operator needs one FE space");
      const int n_elem = sp0_->n_element();
      const int n0 = n_elem ? (int)sp0_->element(0).dof().size() : 0;
      const int n1 = n_elem ? (int)sp1_->element(0).dof().size() : 0;
      std::vector<double> em((size_t)n_elem * n0 * n1);
      for (int e = 0; e < n_elem; e++) {
        getElementPattern(sp0_->element(e), sp1_->element(e));
        element_matrix_.reinit(n0, n1);
        getElementMatrix(sp0_->element(e), sp1_->element(e));
        std::copy(element_matrix_.data(), element_matrix_.data() + (size_t)n0 * n1,
                  em.begin() + (size_t)e * n0 * n1);
      }
      check(afe_assemble_elements(a, em.data()));
    }
    mark_device_newer();
  }
  afe_matrix *pair_dev_ = nullptr; // made by buildSparsityPattern, taken by buildSparseMatrix
  SparsityPattern pattern_store_;
  DenseMatrix<double> element_matrix_;
  int acc_ = 3;
  int n_dof0_ = 0, n_dof1_ = 0, n_max_coupling_ = 0;
  const std::vector<int> *element_dof0_ = nullptr, *element_dof1_ = nullptr;
};

// L2InnerProduct (BilinearOperator.h:142-156, BilinearOperator.templates.h:
// 272-326): the L2 inner product of the basis functions, Jxw * phi0_j *
// phi1_k on the quadrature of element 0.  On one space that is the mass
// matrix's arithmetic, so it runs the device mass kernel; on two spaces of
// one mesh the device two-space kernel.
template <int DIM, class value_type0, class value_type1 = value_type0, int DOW = DIM, int TDIM0 = DIM,
          int TDIM1 = DIM>
class L2InnerProduct : public BilinearOperator<DIM, value_type0, value_type1, DOW, TDIM0, TDIM1> {
  using base = BilinearOperator<DIM, value_type0, value_type1, DOW, TDIM0, TDIM1>;

public:
  L2InnerProduct() = default;
  L2InnerProduct(FEMSpace<value_type0, DIM, DOW, TDIM0> &sp0, FEMSpace<value_type1, DIM, DOW, TDIM1> &sp1)
      : base(sp0, sp1) {}
  virtual ~L2InnerProduct() = default;
  void reinit(FEMSpace<value_type0, DIM, DOW, TDIM0> &sp0, FEMSpace<value_type1, DIM, DOW, TDIM1> &sp1) {
    base::reinit(sp0, sp1);
  }
  // host evaluation, used by classes derived from this one
  virtual void getElementMatrix(const typename base::element0_t &e0, const typename base::element1_t &e1,
                                const typename ActiveElementPairIterator<DIM>::State =
                                    ActiveElementPairIterator<DIM>::EQUAL) {
    const int n0 = (int)e0.dof().size(), n1 = (int)e1.dof().size();
    const double volume = e0.templateElement().volume();
    const auto qi = e0.findQuadratureInfo(this->algebricAccuracy());
    const auto jac = e0.local_to_global_jacobian(qi.quadraturePoint());
    const auto q = e0.local_to_global(qi.quadraturePoint());
    const auto bv0 = e0.basis_function_value(q);
    const auto bv1 = e1.basis_function_value(q);
    for (int l = 0; l < qi.n_quadraturePoint(); l++) {
      const double Jxw = qi.weight(l) * jac[l] * volume;
      for (int j = 0; j < n0; j++)
        for (int k = 0; k < n1; k++) this->elementMatrix(j, k) += Jxw * bv0[j][l] * bv1[k][l];
    }
  }

protected:
  int device_operator() const override { return typeid(*this) == typeid(L2InnerProduct) ? AFE_OP_MASS : -1; }
};

template <int DIM, class value_type, int DOW = DIM, int TDIM = DIM>
class MassMatrix : public BilinearOperator<DIM, value_type, value_type, DOW, TDIM, TDIM> {
  using base = BilinearOperator<DIM, value_type, value_type, DOW, TDIM, TDIM>;

public:
  MassMatrix() = default;
  explicit MassMatrix(FEMSpace<value_type, DIM, DOW, TDIM> &sp) : base(sp, sp) {}
  void reinit(FEMSpace<value_type, DIM, DOW, TDIM> &sp) { base::reinit(sp, sp); }
  // host evaluation (BilinearOperator.templates.h:300-330)
  virtual void getElementMatrix(const Element<value_type, DIM, DOW, TDIM> &e0,
                                const Element<value_type, DIM, DOW, TDIM> &e1,
                                const typename ActiveElementPairIterator<DIM>::State =
                                    ActiveElementPairIterator<DIM>::EQUAL) {
    const int n0 = (int)e0.dof().size(), n1 = (int)e1.dof().size();
    const double volume = e0.templateElement().volume();
    const auto qi = e0.findQuadratureInfo(this->algebricAccuracy());
    const auto jac = e0.local_to_global_jacobian(qi.quadraturePoint());
    const auto q = e0.local_to_global(qi.quadraturePoint());
    const auto bv = e0.basis_function_value(q);
    for (int l = 0; l < qi.n_quadraturePoint(); l++) {
      const double Jxw = qi.weight(l) * jac[l] * volume;
      for (int j = 0; j < n0; j++)
        for (int k = 0; k < n1; k++) this->elementMatrix(j, k) += Jxw * bv[j][l] * bv[k][l];
    }
  }

protected:
  int device_operator() const override { return typeid(*this) == typeid(MassMatrix) ? AFE_OP_MASS : -1; }
};

template <int DIM, class value_type, int DOW = DIM, int TDIM = DIM>
class StiffMatrix : public BilinearOperator<DIM, value_type, value_type, DOW, TDIM, TDIM> {
  using base = BilinearOperator<DIM, value_type, value_type, DOW, TDIM, TDIM>;

public:
  StiffMatrix() = default;
  explicit StiffMatrix(FEMSpace<value_type, DIM, DOW, TDIM> &sp) : base(sp, sp) {}
  void reinit(FEMSpace<value_type, DIM, DOW, TDIM> &sp) { base::reinit(sp, sp); }
  // host evaluation (BilinearOperator.templates.h:374-403)
  virtual void getElementMatrix(const Element<value_type, DIM, DOW, TDIM> &e0,
                                const Element<value_type, DIM, DOW, TDIM> &e1,
                                const typename ActiveElementPairIterator<DIM>::State =
                                    ActiveElementPairIterator<DIM>::EQUAL) {
    const int n0 = (int)e0.dof().size(), n1 = (int)e1.dof().size();
    const double volume = e0.templateElement().volume();
    const auto qi = e0.findQuadratureInfo(this->algebricAccuracy());
    const auto jac = e0.local_to_global_jacobian(qi.quadraturePoint());
    const auto q = e0.local_to_global(qi.quadraturePoint());
    const auto bg = e0.basis_function_gradient(q);
    for (int l = 0; l < qi.n_quadraturePoint(); l++) {
      const double Jxw = qi.weight(l) * jac[l] * volume;
      for (int j = 0; j < n0; j++)
        for (int k = 0; k < n1; k++) this->elementMatrix(j, k) += Jxw * innerProduct(bg[j][l], bg[k][l]);
    }
  }

protected:
  int device_operator() const override { return typeid(*this) == typeid(StiffMatrix) ? AFE_OP_STIFF : -1; }
};

} // namespace AFEPack

using namespace AFEPack;
