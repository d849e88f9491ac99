// Functional.h -- drop-in Functional::L2Error / H1SemiError
// (include/AFEPack/Functional.h:100-110, Functional.templates.h:296-317,
// 399-426 of the reference).  The exact solution is host code, evaluated at
// the device-computed quadrature points; the FE function and the element
// integrals are on the device (afe_error_qp).
#pragma once
#include "FEMSpace.h"

namespace AFEPack {
namespace Functional {

namespace detail_fn {
template <class value_type, int DIM>
double error(FEMFunction<value_type, DIM> &uh, const Function<value_type> &u, int acc, int h1) {
  const auto &sp = uh.femSpace();
  afe_space *h = sp.handle();
  const size_t n_elem = (size_t)sp.n_element();
  int64_t nq = 0;
  AFEPack::detail::check(afe_space_qpoints(h, acc, &nq, nullptr, nullptr));
  std::vector<double> pts(n_elem * nq * DIM);
  AFEPack::detail::check(afe_space_qpoints(h, acc, &nq, pts.data(), nullptr));
  const size_t nqt = n_elem * (size_t)nq;
  std::vector<double> ex(nqt * (h1 ? DIM : 1));
  for (size_t k = 0; k < nqt; k++) {
    if (h1) {
      const std::vector<value_type> g = u.gradient(&pts[k * DIM]);
      if (g.size() < (size_t)DIM) throw std::runtime_error("H1SemiError needs the gradient of the exact solution");
      for (int d = 0; d < DIM; d++) ex[k * DIM + d] = g[d];
    } else {
      ex[k] = u.value(&pts[k * DIM]);
    }
  }
  AFEPack::detail::DeviceVector d(uh.size());
  d.upload(uh.data());
  double err = 0.0;
  AFEPack::detail::check(afe_error_qp(h, d.p, acc, h1, ex.data(), &err));
  return err;
}
} // namespace detail_fn

template <class value_type, int DIM>
value_type L2Error(FEMFunction<value_type, DIM> &uh, const Function<value_type> &u, int acc) {
  return detail_fn::error(uh, u, acc, 0);
}
template <class value_type, int DIM>
value_type H1SemiError(FEMFunction<value_type, DIM> &uh, const Function<value_type> &u, int acc) {
  return detail_fn::error(uh, u, acc, 1);
}

} // namespace Functional
} // namespace AFEPack

using namespace AFEPack;
