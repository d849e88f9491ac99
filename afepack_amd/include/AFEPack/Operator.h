// Operator.h -- drop-in Operator::L2Discretize (include/AFEPack/Operator.h:
// 128-152, Operator.discretize.templates.h:121-190 of the reference).
//
// The right-hand side function is host code: it is evaluated at the physical
// quadrature points the device computed (afe_space_qpoints), and the element
// load vectors and their gather into the DOF vector run on the device
// (afe_rhs_qp).
#pragma once
#include "BilinearOperator.h"
#include "FEMSpace.h"

namespace AFEPack {
namespace Operator {

namespace detail_op {
template <class value_type, int DIM, class F>
void discretize(const FEMSpace<value_type, DIM> &sp, Vector<double> &f1, int acc, const F &f) {
  afe_space *h = sp.handle();
  if (!h) throw std::runtime_error("buildDof() must run before L2Discretize");
  const size_t n_dof = (size_t)sp.n_dof(), n_elem = (size_t)sp.n_element();
  int64_t nq = 0;
  AFEPack::detail::check(afe_space_qpoints(h, acc, &nq, nullptr, nullptr));
  std::vector<double> pts(n_elem * nq * DIM), fq(n_elem * nq);
  AFEPack::detail::check(afe_space_qpoints(h, acc, &nq, pts.data(), nullptr));
  for (size_t k = 0; k < n_elem * (size_t)nq; k++) fq[k] = f(&pts[k * DIM]);
  AFEPack::detail::DeviceVector d(n_dof);
  AFEPack::detail::check(afe_rhs_qp(h, acc, fq.data(), d.p));
  if (f1.size() != n_dof) f1.reinit(n_dof);
  d.download(f1.data());
}
} // namespace detail_op

template <class value_type, int DIM>
void L2Discretize(value_type (*f0)(const double *), const FEMSpace<value_type, DIM> &sp, Vector<double> &f1,
                  int algebric_accuracy) {
  detail_op::discretize(sp, f1, algebric_accuracy, [f0](const double *p) { return (double)f0(p); });
}

template <class value_type, int DIM>
void L2Discretize(const Function<value_type> &f0, const FEMSpace<value_type, DIM> &sp, Vector<double> &f1,
                  int algebric_accuracy) {
  detail_op::discretize(sp, f1, algebric_accuracy, [&f0](const double *p) { return (double)f0.value(p); });
}

} // namespace Operator
} // namespace AFEPack

using namespace AFEPack;
