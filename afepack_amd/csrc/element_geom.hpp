// element_geom.hpp -- template-element plugins of triangle.1, triangle.2 and
// tetrahedron.1, shared by the device kernels and the host-side C++ classes.
//
// Quadrature points are mapped through barycentric area ratios of the
// reference simplex (triangle.crd_trs.c:15-27, tetrahedron.crd_trs.c:18-32);
// basis functions are evaluated at the physical point with the physical
// vertices (triangle.1.bas_fun.c, triangle.2.bas_fun.c,
// tetrahedron.1.bas_fun.c).  Every expression keeps the reference's order of
// operations; compile without floating-point contraction (clang: the pragma
// below; g++: -ffp-contract=off) so that host and device agree bitwise.
#pragma once
#include <cstdint>

#ifdef __HIPCC__
#define AFE_HD __host__ __device__
#else
#define AFE_HD
#endif
#ifdef __clang__
#define AFE_UNROLL _Pragma("unroll")
#else
#define AFE_UNROLL
#endif

namespace afe {

#ifndef AFE_KIND_ENUM
#define AFE_KIND_ENUM
enum { P1_TRI = 1, P2_TRI = 2, P1_TET = 3 };
#endif

#ifdef __clang__
#pragma clang fp contract(off)
#endif

AFE_HD inline double area2(const double *a, const double *b, const double *c) {
  return (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]);
}

AFE_HD inline double det3(const double *v0, const double *v1, const double *v2,
                              const double *v3) {
  const double a0 = v1[0] - v0[0], a1 = v1[1] - v0[1], a2 = v1[2] - v0[2];
  const double b0 = v2[0] - v0[0], b1 = v2[1] - v0[1], b2 = v2[2] - v0[2];
  const double c0 = v3[0] - v0[0], c1 = v3[1] - v0[1], c2 = v3[2] - v0[2];
  return a0 * b1 * c2 + a1 * b2 * c0 + a2 * b0 * c1 - a0 * b2 * c1 - a1 * b0 * c2 - a2 * b1 * c0;
}

template <int KIND> struct KindInfo;
template <> struct KindInfo<P1_TRI> { static constexpr int D = 2, NV = 3, NL = 3; };
template <> struct KindInfo<P2_TRI> { static constexpr int D = 2, NV = 3, NL = 6; };
template <> struct KindInfo<P1_TET> { static constexpr int D = 3, NV = 4, NL = 4; };

template <int KIND>
AFE_HD inline void load_elem(int64_t e, const double *__restrict__ coords,
                                 const int32_t *__restrict__ ev, double v[][KindInfo<KIND>::D]) {
  constexpr int D = KindInfo<KIND>::D, NV = KindInfo<KIND>::NV;
AFE_UNROLL
  for (int j = 0; j < NV; j++) {
    const int64_t p = ev[e * NV + j];
AFE_UNROLL
    for (int d = 0; d < D; d++) v[j][d] = coords[p * D + d];
  }
}

// CoordTransform::local_to_global_jacobian
template <int D> AFE_HD inline double l2g_jacobian(const double v[][D]) {
  if constexpr (D == 2) {
    const double r0[2] = {0.0, 0.0}, r1[2] = {1.0, 0.0}, r2[2] = {0.0, 1.0};
    return area2(v[0], v[1], v[2]) / area2(r0, r1, r2);
  } else {
    const double r0[3] = {0.0, 0.0, 0.0}, r1[3] = {1.0, 0.0, 0.0}, r2[3] = {0.0, 1.0, 0.0},
                 r3[3] = {0.0, 0.0, 1.0};
    return (det3(v[0], v[1], v[2], v[3]) / 6.) / (det3(r0, r1, r2, r3) / 6.);
  }
}

// TemplateGeometry::volume of the reference simplex
template <int D> AFE_HD inline double tmpl_volume() {
  if constexpr (D == 2) {
    return .5 * ((1.0 - 0.0) * (1.0 - 0.0) - (0.0 - 0.0) * (0.0 - 0.0));
  } else {
    const double r0[3] = {0.0, 0.0, 0.0}, r1[3] = {1.0, 0.0, 0.0}, r2[3] = {0.0, 1.0, 0.0},
                 r3[3] = {0.0, 0.0, 1.0};
    return det3(r0, r1, r2, r3) / 6.;
  }
}

// CoordTransform::local_to_global of the reference point lp, in two steps:
// the barycentric coordinates of lp on the reference simplex (area / volume
// ratios; they do not depend on the element, so a kernel may take them from
// a table computed once per quadrature rule), then their combination of the
// physical vertices.  Same operations in the same order as the one-step form.
template <int D> AFE_HD inline void ref_bary(const double *lp, double *lam) {
  if constexpr (D == 2) {
    const double r0[2] = {0.0, 0.0}, r1[2] = {1.0, 0.0}, r2[2] = {0.0, 1.0};
    const double ar = area2(r0, r1, r2);
    lam[0] = area2(lp, r1, r2) / ar;
    lam[1] = area2(lp, r2, r0) / ar;
    lam[2] = area2(lp, r0, r1) / ar;
  } else {
    const double r0[3] = {0.0, 0.0, 0.0}, r1[3] = {1.0, 0.0, 0.0}, r2[3] = {0.0, 1.0, 0.0},
                 r3[3] = {0.0, 0.0, 1.0};
    const double vol = det3(r0, r1, r2, r3) / 6.;
    lam[0] = det3(lp, r1, r2, r3) / 6. / vol;
    lam[1] = det3(r0, lp, r2, r3) / 6. / vol;
    lam[2] = det3(r0, r1, lp, r3) / 6. / vol;
    lam[3] = det3(r0, r1, r2, lp) / 6. / vol;
  }
}

template <int D> AFE_HD inline void bary_to_global(const double *lam, const double v[][D], double *x) {
  if constexpr (D == 2) {
    x[0] = lam[0] * v[0][0] + lam[1] * v[1][0] + lam[2] * v[2][0];
    x[1] = lam[0] * v[0][1] + lam[1] * v[1][1] + lam[2] * v[2][1];
  } else {
AFE_UNROLL
    for (int d = 0; d < 3; d++) x[d] = lam[0] * v[0][d] + lam[1] * v[1][d] + lam[2] * v[2][d] + lam[3] * v[3][d];
  }
}

template <int D> AFE_HD inline void local_to_global(const double *lp, const double v[][D], double *x) {
  double lam[D + 1];
  ref_bary<D>(lp, lam);
  bary_to_global<D>(lam, v, x);
}

// basis values at the physical point p (local order of the .bas_fun files)
template <int KIND>
AFE_HD inline void basis_val(const double *p, const double v[][KindInfo<KIND>::D], double *val) {
  if constexpr (KIND == P1_TRI) {
    const double ar = area2(v[0], v[1], v[2]);
    val[0] = area2(p, v[1], v[2]) / ar;
    val[1] = area2(v[0], p, v[2]) / ar;
    val[2] = area2(v[0], v[1], p) / ar;
  } else if constexpr (KIND == P2_TRI) {
    const double ar = area2(v[0], v[1], v[2]);
    const double l0 = area2(p, v[1], v[2]) / ar, l1 = area2(p, v[2], v[0]) / ar,
                 l2 = area2(p, v[0], v[1]) / ar;
    val[0] = l0 * (2 * l0 - 1.0);
    val[1] = l1 * (2 * l1 - 1.0);
    val[2] = l2 * (2 * l2 - 1.0);
    val[3] = 4.0 * l1 * l2;
    val[4] = 4.0 * l0 * l2;
    val[5] = 4.0 * l0 * l1;
  } else {
    const double vol = det3(v[0], v[1], v[2], v[3]);
    val[0] = det3(p, v[1], v[2], v[3]) / vol;
    val[1] = det3(v[0], p, v[2], v[3]) / vol;
    val[2] = det3(v[0], v[1], p, v[3]) / vol;
    val[3] = det3(v[0], v[1], v[2], p) / vol;
  }
}

// basis gradients at the physical point p
template <int KIND>
AFE_HD inline void basis_grad(const double *p, const double v[][KindInfo<KIND>::D],
                                  double g[][KindInfo<KIND>::D]) {
  if constexpr (KIND == P1_TRI || KIND == P2_TRI) {
    const double ar = area2(v[0], v[1], v[2]);
    double dl[3][2];
AFE_UNROLL
    for (int i = 0; i < 3; i++) {
      const int a = (i + 1) % 3, b = (i + 2) % 3;
      dl[i][0] = (v[a][1] - v[b][1]) / ar;
      dl[i][1] = (v[b][0] - v[a][0]) / ar;
    }
    if constexpr (KIND == P1_TRI) {
AFE_UNROLL
      for (int i = 0; i < 3; i++) {
        g[i][0] = dl[i][0];
        g[i][1] = dl[i][1];
      }
    } else {
      const double l0 = area2(p, v[1], v[2]) / ar, l1 = area2(p, v[2], v[0]) / ar,
                   l2 = area2(p, v[0], v[1]) / ar;
      const double lam[3] = {l0, l1, l2};
AFE_UNROLL
      for (int i = 0; i < 3; i++)
AFE_UNROLL
        for (int d = 0; d < 2; d++) g[i][d] = (4.0 * lam[i] - 1.0) * dl[i][d];
AFE_UNROLL
      for (int d = 0; d < 2; d++) {
        g[3][d] = 4.0 * (l2 * dl[1][d] + l1 * dl[2][d]);
        g[4][d] = 4.0 * (l0 * dl[2][d] + l2 * dl[0][d]);
        g[5][d] = 4.0 * (l1 * dl[0][d] + l0 * dl[1][d]);
      }
    }
  } else {
    const double vol = det3(v[0], v[1], v[2], v[3]);
AFE_UNROLL
    for (int m = 0; m < 4; m++) {
      const int ia = (m + 1) % 4, ib = (m + 2) % 4, ic = (m + 3) % 4;
      const double sgn = (m % 2 == 0) ? -1. : 1.;
AFE_UNROLL
      for (int n = 0; n < 3; n++) {
        const int n1 = (n + 1) % 3, n2 = (n + 2) % 3;
        const double cof = sgn * ((v[ib][n1] - v[ia][n1]) * (v[ic][n2] - v[ia][n2]) -
                                  (v[ib][n2] - v[ia][n2]) * (v[ic][n1] - v[ia][n1]));
        g[m][n] = cof / vol;
      }
    }
  }
}

#ifdef __clang__
#pragma clang fp contract(on)
#endif

} // namespace afe
