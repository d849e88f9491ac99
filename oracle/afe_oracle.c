/*
 * afe_oracle.c -- CPU oracle for AFEPack's assemble-and-solve path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product library.  Each function cites the reference file:line whose
 * algorithm it restates.  Paths are relative to the AFEPack source tree.
 *
 * Floating point: compiled with -ffp-contract=off so that the summation and
 * product order of the reference loops is kept (no fused multiply-add).
 */
#include "afe_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_INVALID INT64_MAX

/* ====================================================================== */
/* quadrature tables                                                       */
/* ====================================================================== */
/* Rules of include/AFEPack/template/triangle/triangle.tmp_geo:37-105 and
 * include/AFEPack/template/tetrahedron/tetrahedron.tmp_geo:63-130, stored
 * with the same decimal strings so that the parsed doubles are identical.
 * Rows are (x, y[, z], weight) on the reference simplex. */
static const double tri_q1[] = {0.3333333333333333, 0.3333333333333333, 1.0};
static const double tri_q2[] = {
    0.166666666666667, 0.166666666666667, 0.333333333333333,
    0.166666666666667, 0.666666666666667, 0.333333333333333,
    0.666666666666667, 0.166666666666667, 0.333333333333333};
static const double tri_q3[] = {
    0.333333333333333, 0.333333333333333, -0.562500000000000,
    0.200000000000000, 0.200000000000000, 0.520833333333333,
    0.200000000000000, 0.600000000000000, 0.520833333333333,
    0.600000000000000, 0.200000000000000, 0.520833333333333};
static const double tri_q4[] = {
    0.445948490915965, 0.445948490915965, 0.223381589678011,
    0.445948490915965, 0.108103018168070, 0.223381589678011,
    0.108103018168070, 0.445948490915965, 0.223381589678011,
    0.091576213509771, 0.091576213509771, 0.109951743655322,
    0.091576213509771, 0.816847572980459, 0.109951743655322,
    0.816847572980459, 0.091576213509771, 0.109951743655322};
static const double tri_q5[] = {
    0.333333333333333, 0.333333333333333, 0.225000000000000,
    0.470142064105115, 0.470142064105115, 0.132394152788506,
    0.470142064105115, 0.059715871789770, 0.132394152788506,
    0.059715871789770, 0.470142064105115, 0.132394152788506,
    0.101286507323456, 0.101286507323456, 0.125939180544827,
    0.101286507323456, 0.797426985353087, 0.125939180544827,
    0.797426985353087, 0.101286507323456, 0.125939180544827};
static const double tri_q6[] = {
    0.249286745170910, 0.249286745170910, 0.116786275726379,
    0.249286745170910, 0.501426509658179, 0.116786275726379,
    0.501426509658179, 0.249286745170910, 0.116786275726379,
    0.063089014491502, 0.063089014491502, 0.050844906370207,
    0.063089014491502, 0.873821971016996, 0.050844906370207,
    0.873821971016996, 0.063089014491502, 0.050844906370207,
    0.310352451033784, 0.636502499121399, 0.082851075618374,
    0.636502499121399, 0.053145049844817, 0.082851075618374,
    0.053145049844817, 0.310352451033784, 0.082851075618374,
    0.636502499121399, 0.310352451033784, 0.082851075618374,
    0.310352451033784, 0.053145049844817, 0.082851075618374,
    0.053145049844817, 0.636502499121399, 0.082851075618374};
static const double tri_q7[] = {
    0.333333333333333, 0.333333333333333, -0.149570044467682,
    0.260345966079040, 0.260345966079040, 0.175615257433208,
    0.260345966079040, 0.479308067841920, 0.175615257433208,
    0.479308067841920, 0.260345966079040, 0.175615257433208,
    0.065130102902216, 0.065130102902216, 0.053347235608838,
    0.065130102902216, 0.869739794195568, 0.053347235608838,
    0.869739794195568, 0.065130102902216, 0.053347235608838,
    0.312865496004874, 0.638444188569810, 0.077113760890257,
    0.638444188569810, 0.048690315425316, 0.077113760890257,
    0.048690315425316, 0.312865496004874, 0.077113760890257,
    0.638444188569810, 0.312865496004874, 0.077113760890257,
    0.312865496004874, 0.048690315425316, 0.077113760890257,
    0.048690315425316, 0.638444188569810, 0.077113760890257};
static const double tet_q1[] = {0.25000000, 0.25000000, 0.25000000, 1.0};
static const double tet_q2[] = {
    0.13819660, 0.13819660, 0.13819660, 0.25,
    0.58541020, 0.13819660, 0.13819660, 0.25,
    0.13819660, 0.58541020, 0.13819660, 0.25,
    0.13819660, 0.13819660, 0.58541020, 0.25};
static const double tet_q3[] = {
    0.25000000, 0.25000000, 0.25000000, -0.8,
    0.16666667, 0.16666667, 0.16666667, 0.45,
    0.50000000, 0.16666667, 0.16666667, 0.45,
    0.16666667, 0.50000000, 0.16666667, 0.45,
    0.16666667, 0.16666667, 0.50000000, 0.45};
static const double tet_q4[] = {
    0.2500000000000000, 0.2500000000000000, 0.2500000000000000, -0.0789333333333333,
    0.7857142857142857, 0.0714285714285714, 0.0714285714285714, 0.0457333333333333,
    0.0714285714285714, 0.0714285714285714, 0.0714285714285714, 0.0457333333333333,
    0.0714285714285714, 0.0714285714285714, 0.7857142857142857, 0.0457333333333333,
    0.0714285714285714, 0.7857142857142857, 0.0714285714285714, 0.0457333333333333,
    0.1005964238332008, 0.3994035761667992, 0.3994035761667992, 0.1493333333333333,
    0.3994035761667992, 0.1005964238332008, 0.3994035761667992, 0.1493333333333333,
    0.3994035761667992, 0.3994035761667992, 0.1005964238332008, 0.1493333333333333,
    0.3994035761667992, 0.1005964238332008, 0.1005964238332008, 0.1493333333333333,
    0.1005964238332008, 0.3994035761667992, 0.1005964238332008, 0.1493333333333333,
    0.1005964238332008, 0.1005964238332008, 0.3994035761667992, 0.1493333333333333};
static const double tet_q6[] = {
    0.3561913862225449, 0.2146028712591517, 0.2146028712591517, 0.0399227502581679,
    0.2146028712591517, 0.2146028712591517, 0.2146028712591517, 0.0399227502581679,
    0.2146028712591517, 0.2146028712591517, 0.3561913862225449, 0.0399227502581679,
    0.2146028712591517, 0.3561913862225449, 0.2146028712591517, 0.0399227502581679,
    0.8779781243961660, 0.0406739585346113, 0.0406739585346113, 0.0100772110553207,
    0.0406739585346113, 0.0406739585346113, 0.0406739585346113, 0.0100772110553207,
    0.0406739585346113, 0.0406739585346113, 0.8779781243961660, 0.0100772110553207,
    0.0406739585346113, 0.8779781243961660, 0.0406739585346113, 0.0100772110553207,
    0.0329863295731731, 0.3223378901422757, 0.3223378901422757, 0.0553571815436544,
    0.3223378901422757, 0.3223378901422757, 0.3223378901422757, 0.0553571815436544,
    0.3223378901422757, 0.3223378901422757, 0.0329863295731731, 0.0553571815436544,
    0.3223378901422757, 0.0329863295731731, 0.3223378901422757, 0.0553571815436544,
    0.2696723314583159, 0.0636610018750175, 0.0636610018750175, 0.0482142857142857,
    0.0636610018750175, 0.2696723314583159, 0.0636610018750175, 0.0482142857142857,
    0.0636610018750175, 0.0636610018750175, 0.2696723314583159, 0.0482142857142857,
    0.6030056647916491, 0.0636610018750175, 0.0636610018750175, 0.0482142857142857,
    0.0636610018750175, 0.6030056647916491, 0.0636610018750175, 0.0482142857142857,
    0.0636610018750175, 0.0636610018750175, 0.6030056647916491, 0.0482142857142857,
    0.0636610018750175, 0.2696723314583159, 0.6030056647916491, 0.0482142857142857,
    0.2696723314583159, 0.6030056647916491, 0.0636610018750175, 0.0482142857142857,
    0.6030056647916491, 0.0636610018750175, 0.2696723314583159, 0.0482142857142857,
    0.0636610018750175, 0.6030056647916491, 0.2696723314583159, 0.0482142857142857,
    0.2696723314583159, 0.0636610018750175, 0.6030056647916491, 0.0482142857142857,
    0.6030056647916491, 0.2696723314583159, 0.0636610018750175, 0.0482142857142857};

typedef struct { int acc, n; const double *d; } qrule;
static const qrule tri_rules[] = {{1, 1, tri_q1}, {2, 3, tri_q2}, {3, 4, tri_q3},
                                  {4, 6, tri_q4}, {5, 7, tri_q5}, {6, 12, tri_q6},
                                  {7, 13, tri_q7}};
static const qrule tet_rules[] = {{1, 1, tet_q1}, {2, 4, tet_q2}, {3, 5, tet_q3},
                                  {4, 11, tet_q4}, {6, 24, tet_q6}};

/* QuadratureInfoAdmin::find: the first rule whose algebraic accuracy is
 * >= acc (include/AFEPack/Geometry.templates.h:672-704). */
static const qrule *quad_find(int tdim, int acc) {
  const qrule *r = tdim == 2 ? tri_rules : tet_rules;
  int n = tdim == 2 ? (int)(sizeof tri_rules / sizeof *r)
                    : (int)(sizeof tet_rules / sizeof *r);
  for (int i = 0; i < n; i++)
    if (r[i].acc >= acc) return &r[i];
  return NULL;
}

int or_quad_n(int tdim, int acc) {
  const qrule *q = quad_find(tdim, acc);
  return q ? q->n : -1;
}

void or_quad_get(int tdim, int acc, double *pts, double *w) {
  const qrule *q = quad_find(tdim, acc);
  if (!q) return;
  for (int l = 0; l < q->n; l++) {
    for (int d = 0; d < tdim; d++) pts[l * tdim + d] = q->d[l * (tdim + 1) + d];
    w[l] = q->d[l * (tdim + 1) + tdim];
  }
}

/* ====================================================================== */
/* template-element geometry and basis functions                           */
/* ====================================================================== */
/* signed doubled area, template/triangle/triangle.crd_trs.c:11-12 */
static double area2(const double *a, const double *b, const double *c) {
  return (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]);
}

/* 6 x signed volume, template/tetrahedron/tetrahedron.1/tetrahedron.1.bas_fun.c:5-16 */
static double det3(const double *v0, const double *v1, const double *v2,
                   const double *v3) {
  double a0 = v1[0] - v0[0], a1 = v1[1] - v0[1], a2 = v1[2] - v0[2];
  double b0 = v2[0] - v0[0], b1 = v2[1] - v0[1], b2 = v2[2] - v0[2];
  double c0 = v3[0] - v0[0], c1 = v3[1] - v0[1], c2 = v3[2] - v0[2];
  return a0 * b1 * c2 + a1 * b2 * c0 + a2 * b0 * c1 - a0 * b2 * c1 -
         a1 * b0 * c2 - a2 * b1 * c0;
}

static const double tri_ref[3][2] = {{0.0, 0.0}, {1.0, 0.0}, {0.0, 1.0}};
static const double tet_ref[4][3] = {
    {0.0, 0.0, 0.0}, {1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};

/* template volume: triangle.volume.c / tetrahedron.volume.c evaluated on the
 * reference vertices (TemplateGeometry::volume, Geometry.templates.h:818-830) */
static double tmpl_volume(int tdim) {
  if (tdim == 2)
    return .5 * ((tri_ref[1][0] - tri_ref[0][0]) * (tri_ref[2][1] - tri_ref[0][1]) -
                 (tri_ref[1][1] - tri_ref[0][1]) * (tri_ref[2][0] - tri_ref[0][0]));
  return det3(tet_ref[0], tet_ref[1], tet_ref[2], tet_ref[3]) / 6.;
}

/* CoordTransform::local_to_global (triangle.crd_trs.c:15-27,
 * tetrahedron.crd_trs.c:18-32): barycentric coordinates from the reference
 * simplex, then the same combination of the physical vertices. */
static void local_to_global(int tdim, const double *lp, const double *gv,
                            double *gp) {
  if (tdim == 2) {
    double ar = area2(tri_ref[0], tri_ref[1], tri_ref[2]);
    double l0 = area2(lp, tri_ref[1], tri_ref[2]) / ar;
    double l1 = area2(lp, tri_ref[2], tri_ref[0]) / ar;
    double l2 = area2(lp, tri_ref[0], tri_ref[1]) / ar;
    gp[0] = l0 * gv[0] + l1 * gv[2] + l2 * gv[4];
    gp[1] = l0 * gv[1] + l1 * gv[3] + l2 * gv[5];
  } else {
    double vol = det3(tet_ref[0], tet_ref[1], tet_ref[2], tet_ref[3]) / 6.;
    double l0 = det3(lp, tet_ref[1], tet_ref[2], tet_ref[3]) / 6. / vol;
    double l1 = det3(tet_ref[0], lp, tet_ref[2], tet_ref[3]) / 6. / vol;
    double l2 = det3(tet_ref[0], tet_ref[1], lp, tet_ref[3]) / 6. / vol;
    double l3 = det3(tet_ref[0], tet_ref[1], tet_ref[2], lp) / 6. / vol;
    for (int d = 0; d < 3; d++)
      gp[d] = l0 * gv[d] + l1 * gv[3 + d] + l2 * gv[6 + d] + l3 * gv[9 + d];
  }
}

/* CoordTransform::local_to_global_jacobian: ratio of the physical to the
 * reference signed measure (triangle.crd_trs.c:43-50,
 * tetrahedron.crd_trs.c:50-57). */
static double l2g_jacobian(int tdim, const double *gv) {
  if (tdim == 2)
    return area2(gv, gv + 2, gv + 4) / area2(tri_ref[0], tri_ref[1], tri_ref[2]);
  return (det3(gv, gv + 3, gv + 6, gv + 9) / 6.) /
         (det3(tet_ref[0], tet_ref[1], tet_ref[2], tet_ref[3]) / 6.);
}

/* barycentric coordinates of a physical point p in the physical triangle
 * (triangle.2.bas_fun.c:21-27) */
static void tri_lambda(const double *p, const double *v, double *lam,
                       double *ar) {
  *ar = area2(v, v + 2, v + 4);
  lam[0] = area2(p, v + 2, v + 4) / *ar;
  lam[1] = area2(p, v + 4, v) / *ar;
  lam[2] = area2(p, v, v + 2) / *ar;
}

/* Basis values at a physical point with physical vertices.
 * P1 triangle: triangle.1/triangle.1.bas_fun.c:19-41 (lambda_i);
 * P2 triangle: triangle.2/triangle.2.bas_fun.c:29-75, local order = vertices
 *   0,1,2 then edges 0,1,2 (edge k opposite vertex k, triangle.2.bas_fun:1-37);
 * P1 tet: tetrahedron.1/tetrahedron.1.bas_fun.c:28-58. */
static void basis_value(int kind, const double *p, const double *v,
                        double *val) {
  if (kind == OR_P1_TRI) {
    double ar = area2(v, v + 2, v + 4);
    val[0] = area2(p, v + 2, v + 4) / ar;
    val[1] = area2(v, p, v + 4) / ar;
    val[2] = area2(v, v + 2, p) / ar;
  } else if (kind == OR_P2_TRI) {
    double l[3], ar;
    tri_lambda(p, v, l, &ar);
    val[0] = l[0] * (2 * l[0] - 1.0);
    val[1] = l[1] * (2 * l[1] - 1.0);
    val[2] = l[2] * (2 * l[2] - 1.0);
    val[3] = 4.0 * l[1] * l[2];
    val[4] = 4.0 * l[0] * l[2];
    val[5] = 4.0 * l[0] * l[1];
  } else {
    double vol = det3(v, v + 3, v + 6, v + 9);
    val[0] = det3(p, v + 3, v + 6, v + 9) / vol;
    val[1] = det3(v, p, v + 6, v + 9) / vol;
    val[2] = det3(v, v + 3, p, v + 9) / vol;
    val[3] = det3(v, v + 3, v + 6, p) / vol;
  }
}

/* Basis gradients at a physical point, layout g[basis*dim + d].
 * P1 triangle: triangle.1.bas_fun.c:13-17,43-62 (scaled rotated opposite edge);
 * P2 triangle: triangle.2.bas_fun.c:13-18,77-129;
 * P1 tet: tetrahedron.1.bas_fun.c:19-26,60-95 -- the gradient of lambda_m is
 *   the cross product of the two edges of the opposite face spanned from
 *   vertex (m+1)%4, negated for even m, over the 6x volume. */
static void basis_grad(int kind, const double *p, const double *v, double *g) {
  if (kind == OR_P1_TRI) {
    double ar = area2(v, v + 2, v + 4);
    for (int i = 0; i < 3; i++) {
      const double *a = v + 2 * ((i + 1) % 3), *b = v + 2 * ((i + 2) % 3);
      g[2 * i + 0] = (a[1] - b[1]) / ar;
      g[2 * i + 1] = (b[0] - a[0]) / ar;
    }
  } else if (kind == OR_P2_TRI) {
    double l[3], ar;
    tri_lambda(p, v, l, &ar);
    double dl[3][2];
    for (int i = 0; i < 3; i++) {
      const double *a = v + 2 * ((i + 1) % 3), *b = v + 2 * ((i + 2) % 3);
      dl[i][0] = (a[1] - b[1]) / ar;
      dl[i][1] = (b[0] - a[0]) / ar;
    }
    for (int i = 0; i < 3; i++)
      for (int d = 0; d < 2; d++) g[2 * i + d] = (4.0 * l[i] - 1.0) * dl[i][d];
    for (int d = 0; d < 2; d++) {
      g[6 + d] = 4.0 * (l[2] * dl[1][d] + l[1] * dl[2][d]);
      g[8 + d] = 4.0 * (l[0] * dl[2][d] + l[2] * dl[0][d]);
      g[10 + d] = 4.0 * (l[1] * dl[0][d] + l[0] * dl[1][d]);
    }
  } else {
    double vol = det3(v, v + 3, v + 6, v + 9);
    for (int m = 0; m < 4; m++) {
      const double *a = v + 3 * ((m + 1) % 4), *b = v + 3 * ((m + 2) % 4),
                   *c = v + 3 * ((m + 3) % 4);
      double sgn = (m % 2 == 0) ? -1. : 1.;
      for (int n = 0; n < 3; n++) {
        int n1 = (n + 1) % 3, n2 = (n + 2) % 3;
        double cof = sgn * ((b[n1] - a[n1]) * (c[n2] - a[n2]) -
                            (b[n2] - a[n2]) * (c[n1] - a[n1]));
        g[3 * m + n] = cof / vol;
      }
    }
  }
}

static int kind_tdim(int kind) { return kind == OR_P1_TET ? 3 : 2; }

void or_basis_eval(int kind, const double *p, const double *verts, double *val,
                   double *grad) {
  if (val) basis_value(kind, p, verts, val);
  if (grad) basis_grad(kind, p, verts, grad);
}
void or_local_to_global(int tdim, const double *lp, const double *verts, double *gp) {
  local_to_global(tdim, lp, verts, gp);
}
double or_l2g_jacobian(int tdim, const double *verts) { return l2g_jacobian(tdim, verts); }
double or_template_volume(int tdim) { return tmpl_volume(tdim); }
static int kind_nloc(int kind) { return kind == OR_P1_TRI ? 3 : kind == OR_P2_TRI ? 6 : 4; }
static int kind_nvtx(int kind) { return kind == OR_P1_TET ? 4 : 3; }

/* coefficient a(x) = exp(g(x)) of the variable-coefficient operators, and
 * grad g(x) (used for the manufactured right-hand side) */
static double field_g(const or_field *f, const double *x, double *grad) {
  double g = 0.0;
  if (grad)
    for (int d = 0; d < f->dim; d++) grad[d] = 0.0;
  for (int m = 0; m < f->n_modes; m++) {
    double arg = f->phi[m];
    for (int d = 0; d < f->dim; d++) arg += 2.0 * M_PI * f->k[m * f->dim + d] * x[d];
    g += f->w[m] * sin(arg);
    if (grad) {
      double c = f->w[m] * cos(arg);
      for (int d = 0; d < f->dim; d++) grad[d] += c * 2.0 * M_PI * f->k[m * f->dim + d];
    }
  }
  if (grad)
    for (int d = 0; d < f->dim; d++) grad[d] *= f->scale;
  return f->scale * g;
}

/* ====================================================================== */
/* FE space: DOF numbering                                                 */
/* ====================================================================== */
struct or_space {
  int kind, tdim, n_loc, n_vtx;
  int64_t n_vert, n_elem, n_dof;
  double *coords;
  int32_t *elem_vert;
  int32_t *elem_dof;
  int32_t *dof_mark;
  double *dof_point;
};

typedef struct { int64_t a, b, e; } edge_rec;
static int edge_cmp(const void *x, const void *y) {
  const edge_rec *p = x, *q = y;
  if (p->a != q->a) return p->a < q->a ? -1 : 1;
  if (p->b != q->b) return p->b < q->b ? -1 : 1;
  return p->e < q->e ? -1 : (p->e > q->e);
}

/* FEMSpace::buildDof, single-thread order (include/AFEPack/FEMSpace.templates.h:
 * 705-880 with n_thread = 1, src/Thread.cpp:13): elements in order, then
 * dimensions 0..TDIM, then the template sub-geometries of that dimension in
 * template order; a geometry gets its DOFs on first touch.  Triangle edge k is
 * the edge opposite vertex k (triangle.tmp_geo:21-30), matched to the mesh
 * edge through the vertex set (FEMSpace.templates.h:183-258).
 * DOF boundary marks follow buildDofBoundaryMark (:689-701): a vertex DOF
 * takes the vertex mark, an edge DOF the mark of its edge, which is 1 for an
 * edge of a single triangle (a boundary side) and 0 otherwise. */
or_space *or_space_new(int kind, int64_t n_vert, const double *coords,
                       int64_t n_elem, const int32_t *elem_vert,
                       const int32_t *vert_mark) {
  or_space *s = calloc(1, sizeof *s);
  s->kind = kind;
  s->tdim = kind_tdim(kind);
  s->n_loc = kind_nloc(kind);
  s->n_vtx = kind_nvtx(kind);
  s->n_vert = n_vert;
  s->n_elem = n_elem;
  int dim = s->tdim, nv = s->n_vtx, nl = s->n_loc;
  s->coords = malloc(sizeof(double) * n_vert * dim);
  memcpy(s->coords, coords, sizeof(double) * n_vert * dim);
  s->elem_vert = malloc(sizeof(int32_t) * n_elem * nv);
  memcpy(s->elem_vert, elem_vert, sizeof(int32_t) * n_elem * nv);
  s->elem_dof = malloc(sizeof(int32_t) * n_elem * nl);

  int64_t *vdof = malloc(sizeof(int64_t) * n_vert);
  for (int64_t i = 0; i < n_vert; i++) vdof[i] = -1;
  int64_t n_edge = 0, *edge_of = NULL, *edof = NULL;
  int32_t *edge_cnt = NULL;
  if (kind == OR_P2_TRI) {
    /* mesh edges: unordered vertex pairs, one record per (element, edge) */
    edge_rec *er = malloc(sizeof(edge_rec) * n_elem * 3);
    for (int64_t e = 0; e < n_elem; e++)
      for (int k = 0; k < 3; k++) {
        int64_t p = elem_vert[e * 3 + (k + 1) % 3], q = elem_vert[e * 3 + (k + 2) % 3];
        er[e * 3 + k].a = p < q ? p : q;
        er[e * 3 + k].b = p < q ? q : p;
        er[e * 3 + k].e = e * 3 + k;
      }
    qsort(er, n_elem * 3, sizeof *er, edge_cmp);
    edge_of = malloc(sizeof(int64_t) * n_elem * 3);
    edge_cnt = calloc(n_elem * 3, sizeof(int32_t));
    for (int64_t i = 0; i < n_elem * 3; i++) {
      if (i > 0 && (er[i].a != er[i - 1].a || er[i].b != er[i - 1].b)) n_edge++;
      edge_of[er[i].e] = n_edge;
      edge_cnt[n_edge]++;
    }
    n_edge++;
    free(er);
    edof = malloc(sizeof(int64_t) * n_edge);
    for (int64_t i = 0; i < n_edge; i++) edof[i] = -1;
  }
  int64_t n_dof = 0;
  for (int64_t e = 0; e < n_elem; e++) {
    for (int j = 0; j < nv; j++) {
      int64_t v = elem_vert[e * nv + j];
      if (vdof[v] < 0) vdof[v] = n_dof++;
      s->elem_dof[e * nl + j] = (int32_t)vdof[v];
    }
    if (kind == OR_P2_TRI)
      for (int j = 0; j < 3; j++) {
        int64_t g = edge_of[e * 3 + j];
        if (edof[g] < 0) edof[g] = n_dof++;
        s->elem_dof[e * nl + 3 + j] = (int32_t)edof[g];
      }
  }
  s->n_dof = n_dof;
  s->dof_mark = calloc(n_dof, sizeof(int32_t));
  s->dof_point = calloc(n_dof * dim, sizeof(double));
  for (int64_t e = 0; e < n_elem; e++) {
    double gv[12];
    for (int j = 0; j < nv; j++)
      for (int d = 0; d < dim; d++)
        gv[j * dim + d] = coords[(int64_t)elem_vert[e * nv + j] * dim + d];
    for (int j = 0; j < nv; j++) {
      int64_t dof = s->elem_dof[e * nl + j];
      s->dof_mark[dof] = vert_mark[elem_vert[e * nv + j]];
      for (int d = 0; d < dim; d++) s->dof_point[dof * dim + d] = gv[j * dim + d];
    }
    if (kind == OR_P2_TRI) {
      /* interpolation points of the edge DOFs: (0.5,0.5), (0,0.5), (0.5,0)
       * mapped by local_to_global (triangle.2.bas_fun:19-37) */
      static const double ip[3][2] = {{0.5, 0.5}, {0.0, 0.5}, {0.5, 0.0}};
      for (int j = 0; j < 3; j++) {
        int64_t dof = s->elem_dof[e * nl + 3 + j];
        s->dof_mark[dof] = edge_cnt[edge_of[e * 3 + j]] == 1 ? 1 : 0;
        local_to_global(2, ip[j], gv, s->dof_point + dof * 2);
      }
    }
  }
  free(vdof);
  free(edge_of);
  free(edof);
  free(edge_cnt);
  return s;
}

/* or_space_new with the mesh's side list (Mesh geometry(1): vertex pairs and
 * boundary marks, as EasyMesh .s files give them): buildDofBoundaryMark
 * (FEMSpace.templates.h:689-701) gives an edge DOF the boundary mark of its
 * side.  NULL when an element edge is not among the sides. */
static int64_t side_key(int64_t a, int64_t b) { return a < b ? (a << 32) | b : (b << 32) | a; }
typedef struct {
  int64_t key;
  int32_t mark;
} side_rec;
static int side_cmp(const void *x, const void *y) {
  const side_rec *a = x, *b = y;
  return a->key < b->key ? -1 : a->key > b->key;
}
or_space *or_space_new_sides(int kind, int64_t n_vert, const double *coords,
                             int64_t n_elem, const int32_t *elem_vert,
                             const int32_t *vert_mark, int64_t n_side,
                             const int32_t *side_vert, const int32_t *side_mark) {
  or_space *s = or_space_new(kind, n_vert, coords, n_elem, elem_vert, vert_mark);
  if (kind != OR_P2_TRI || n_side <= 0) return s;
  side_rec *sr = malloc(sizeof(side_rec) * n_side);
  for (int64_t i = 0; i < n_side; i++) {
    sr[i].key = side_key(side_vert[2 * i], side_vert[2 * i + 1]);
    sr[i].mark = side_mark[i];
  }
  qsort(sr, n_side, sizeof *sr, side_cmp);
  for (int64_t e = 0; e < n_elem; e++)
    for (int j = 0; j < 3; j++) {
      side_rec k = {side_key(elem_vert[e * 3 + (j + 1) % 3], elem_vert[e * 3 + (j + 2) % 3]), 0};
      const side_rec *hit = bsearch(&k, sr, n_side, sizeof *sr, side_cmp);
      if (!hit) {
        free(sr);
        or_space_free(s);
        return NULL;
      }
      s->dof_mark[s->elem_dof[e * 6 + 3 + j]] = hit->mark;
    }
  free(sr);
  return s;
}

void or_space_free(or_space *s) {
  if (!s) return;
  free(s->coords);
  free(s->elem_vert);
  free(s->elem_dof);
  free(s->dof_mark);
  free(s->dof_point);
  free(s);
}

int64_t or_space_n_dof(const or_space *s) { return s->n_dof; }
int or_space_n_loc(const or_space *s) { return s->n_loc; }

void or_space_dofs(const or_space *s, int32_t *elem_dof, int32_t *dof_mark,
                   double *dof_point) {
  if (elem_dof) memcpy(elem_dof, s->elem_dof, sizeof(int32_t) * s->n_elem * s->n_loc);
  if (dof_mark) memcpy(dof_mark, s->dof_mark, sizeof(int32_t) * s->n_dof);
  if (dof_point) memcpy(dof_point, s->dof_point, sizeof(double) * s->n_dof * s->tdim);
}

static void elem_coords(const or_space *s, int64_t e, double *gv) {
  int nv = s->n_vtx, dim = s->tdim;
  for (int j = 0; j < nv; j++)
    for (int d = 0; d < dim; d++)
      gv[j * dim + d] = s->coords[(int64_t)s->elem_vert[e * nv + j] * dim + d];
}

/* ====================================================================== */
/* SparsityPattern / SparseMatrix                                          */
/* ====================================================================== */
typedef struct or_sp {
  int64_t m, n, nnz, maxlen;
  int diag_first;
  int compressed;
  int refs;
  int64_t *rp, *col;
} or_sp;

struct or_mat {
  or_sp *sp;
  double *val;
};

/* SparsityPattern::reinit(m, n, row_lengths), SparsityPattern.h:16-74 */
static or_sp *sp_new_lengths(int64_t m, int64_t n, const int64_t *len) {
  or_sp *sp = calloc(1, sizeof *sp);
  sp->m = m;
  sp->n = n;
  sp->refs = 1;
  sp->diag_first = (m == n);
  sp->rp = malloc(sizeof(int64_t) * (m + 1));
  sp->rp[0] = 0;
  for (int64_t i = 0; i < m; i++) {
    sp->rp[i + 1] = sp->rp[i] + len[i];
    if (len[i] > sp->maxlen) sp->maxlen = len[i];
  }
  sp->nnz = sp->rp[m];
  sp->col = malloc(sizeof(int64_t) * (sp->nnz > 0 ? sp->nnz : 1));
  for (int64_t k = 0; k < sp->nnz; k++) sp->col[k] = OR_INVALID;
  if (sp->diag_first)
    for (int64_t i = 0; i < m; i++) sp->col[sp->rp[i]] = i;
  return sp;
}

/* SparsityPattern::reinit(m, n, max_row_length), SparsityPattern.h:76-126 */
static or_sp *sp_new_max(int64_t m, int64_t n, int64_t maxlen) {
  int64_t *len = malloc(sizeof(int64_t) * (m > 0 ? m : 1));
  for (int64_t i = 0; i < m; i++) len[i] = maxlen;
  or_sp *sp = sp_new_lengths(m, n, len);
  free(len);
  return sp;
}

/* SparsityPattern::add, SparsityPattern.h:251-262: first free slot of the
 * row, silently dropped when the row is full */
static void sp_add(or_sp *sp, int64_t i, int64_t j) {
  for (int64_t k = sp->rp[i]; k < sp->rp[i + 1]; k++) {
    if (sp->col[k] == j) return;
    if (sp->col[k] == OR_INVALID) {
      sp->col[k] = j;
      return;
    }
  }
}

static int i64_cmp(const void *x, const void *y) {
  int64_t a = *(const int64_t *)x, b = *(const int64_t *)y;
  return a < b ? -1 : (a > b);
}

/* SparsityPattern::compress, SparsityPattern.h:193-237: drop the free slots;
 * sort each row, except the leading diagonal of a square pattern */
static void sp_compress(or_sp *sp) {
  if (sp->compressed) return;
  int64_t w = 0;
  int64_t start = 0;
  for (int64_t i = 0; i < sp->m; i++) {
    int64_t end = sp->rp[i + 1], cnt = 0;
    for (int64_t k = start; k < end; k++) {
      if (sp->col[k] == OR_INVALID) break;
      sp->col[w + cnt] = sp->col[k];
      cnt++;
    }
    int64_t s0 = sp->diag_first ? 1 : 0;
    if (cnt > s0) qsort(sp->col + w + s0, cnt - s0, sizeof(int64_t), i64_cmp);
    sp->rp[i] = w;
    w += cnt;
    start = end;
  }
  sp->rp[sp->m] = w;
  sp->nnz = w;
  sp->compressed = 1;
}

/* SparsityPattern::max_entries_per_row, SparsityPattern.h:239-249 (the loop
 * starts at row 1 and compares row i-1, so the last row is not visited) */
static int64_t sp_max_entries_per_row(const or_sp *sp) {
  if (!sp->compressed) return sp->maxlen;
  int64_t m = 0;
  for (int64_t i = 1; i < sp->m; i++)
    if (sp->rp[i] - sp->rp[i - 1] > m) m = sp->rp[i] - sp->rp[i - 1];
  return m;
}

static void sp_release(or_sp *sp) {
  if (sp && --sp->refs == 0) {
    free(sp->rp);
    free(sp->col);
    free(sp);
  }
}

static or_mat *mat_new(or_sp *sp) {
  or_mat *a = calloc(1, sizeof *a);
  a->sp = sp;
  a->val = calloc(sp->nnz > 0 ? sp->nnz : 1, sizeof(double));
  return a;
}

/* SparseMatrix::add, SparseMatrix.h:35-50 (row scan; out-of-pattern dropped) */
static void mat_add(or_mat *a, int64_t i, int64_t j, double v) {
  const int64_t *rp = a->sp->rp, *col = a->sp->col;
  for (int64_t k = rp[i]; k < rp[i + 1]; k++)
    if (col[k] == j) {
      a->val[k] += v;
      return;
    }
}

/* SparseMatrix::el, SparseMatrix.h:58-70 */
static double mat_el(const or_mat *a, int64_t i, int64_t j) {
  const int64_t *rp = a->sp->rp, *col = a->sp->col;
  for (int64_t k = rp[i]; k < rp[i + 1]; k++)
    if (col[k] == j) return a->val[k];
  return 0;
}

or_mat *or_mat_from_csr(int64_t m, int64_t n, const int64_t *rp,
                        const int64_t *col, const double *val) {
  or_sp *sp = calloc(1, sizeof *sp);
  sp->m = m;
  sp->n = n;
  sp->refs = 1;
  sp->diag_first = (m == n);
  sp->compressed = 1;
  sp->nnz = rp[m];
  sp->rp = malloc(sizeof(int64_t) * (m + 1));
  memcpy(sp->rp, rp, sizeof(int64_t) * (m + 1));
  sp->col = malloc(sizeof(int64_t) * (sp->nnz > 0 ? sp->nnz : 1));
  memcpy(sp->col, col, sizeof(int64_t) * sp->nnz);
  for (int64_t i = 0; i < m; i++)
    if (rp[i + 1] - rp[i] > sp->maxlen) sp->maxlen = rp[i + 1] - rp[i];
  or_mat *a = mat_new(sp);
  if (val) memcpy(a->val, val, sizeof(double) * sp->nnz);
  return a;
}

/* a matrix on the same SparsityPattern object (the object identity that
 * AMGSolver::lazyReinit checks) */
or_mat *or_mat_share_pattern(const or_mat *a, const double *val) {
  a->sp->refs++;
  or_mat *b = mat_new(a->sp);
  if (val) memcpy(b->val, val, sizeof(double) * a->sp->nnz);
  return b;
}

void or_mat_free(or_mat *a) {
  if (!a) return;
  sp_release(a->sp);
  free(a->val);
  free(a);
}

int64_t or_mat_m(const or_mat *a) { return a->sp->m; }
int64_t or_mat_n(const or_mat *a) { return a->sp->n; }
int64_t or_mat_nnz(const or_mat *a) { return a->sp->nnz; }

void or_mat_get(const or_mat *a, int64_t *rp, int64_t *col, double *val) {
  if (rp) memcpy(rp, a->sp->rp, sizeof(int64_t) * (a->sp->m + 1));
  if (col) memcpy(col, a->sp->col, sizeof(int64_t) * a->sp->nnz);
  if (val) memcpy(val, a->val, sizeof(double) * a->sp->nnz);
}

/* SparseMatrix::vmult, SparseMatrix.h:100-114 */
void or_mat_vmult(const or_mat *a, const double *x, double *y) {
  const int64_t *rp = a->sp->rp, *col = a->sp->col;
  for (int64_t i = 0; i < a->sp->m; i++) {
    double s = 0;
    for (int64_t k = rp[i]; k < rp[i + 1]; k++) s += a->val[k] * x[col[k]];
    y[i] = s;
  }
}

/* SparseMatrix::Tvmult, SparseMatrix.h:116-132 */
void or_mat_tvmult(const or_mat *a, const double *x, double *y) {
  const int64_t *rp = a->sp->rp, *col = a->sp->col;
  for (int64_t j = 0; j < a->sp->n; j++) y[j] = 0;
  for (int64_t i = 0; i < a->sp->m; i++)
    for (int64_t k = rp[i]; k < rp[i + 1]; k++) y[col[k]] += a->val[k] * x[i];
}

/* ====================================================================== */
/* element matrices and assembly                                           */
/* ====================================================================== */
/* StiffMatrix::getElementMatrix (BilinearOperator.templates.h:382-406),
 * MassMatrix::getElementMatrix (:342-366) and the variable-coefficient
 * override of examples/elliptic_equation/EllipticEquation.cpp:65-108:
 *   Jxw = w_l * jacobian_l * volume;  K(j,k) += Jxw [* a(x_l)] * <g_j, g_k>.
 * Quadrature points are mapped to the physical element first and the basis
 * plugins are evaluated there with the physical vertices. */
void or_element_matrix(int kind, int op, int acc, const double *gv,
                       const or_field *coef, double *K) {
  int tdim = kind_tdim(kind), nl = kind_nloc(kind);
  const qrule *q = quad_find(tdim, acc);
  double volume = tmpl_volume(tdim);
  double jac = l2g_jacobian(tdim, gv);
  for (int i = 0; i < nl * nl; i++) K[i] = 0.0;
  for (int l = 0; l < q->n; l++) {
    double lp[3], xp[3], g[18], v[6];
    for (int d = 0; d < tdim; d++) lp[d] = q->d[l * (tdim + 1) + d];
    double w = q->d[l * (tdim + 1) + tdim];
    local_to_global(tdim, lp, gv, xp);
    double Jxw = w * jac * volume;
    if (op == OR_OP_MASS) {
      basis_value(kind, xp, gv, v);
      for (int j = 0; j < nl; j++)
        for (int k = 0; k < nl; k++) K[j * nl + k] += Jxw * v[j] * v[k];
    } else {
      basis_grad(kind, xp, gv, g);
      double a = op == OR_OP_VARSTIFF ? exp(field_g(coef, xp, NULL)) : 1.0;
      for (int j = 0; j < nl; j++)
        for (int k = 0; k < nl; k++) {
          double ip = 0;
          for (int d = 0; d < tdim; d++) ip += g[j * tdim + d] * g[k * tdim + d];
          if (op == OR_OP_VARSTIFF)
            K[j * nl + k] += Jxw * a * ip;
          else
            K[j * nl + k] += Jxw * ip;
        }
    }
  }
}

/* BilinearOperator::build = buildSparsityPattern + buildSparseMatrix
 * (BilinearOperator.templates.h:55-111,115-172,176-255):
 *  - max coupling count = max over DOFs of the summed element DOF counts,
 *    capped at the column count;
 *  - reinit(N, N, max coupling) puts the diagonal first; each element adds
 *    all its (i, j) pairs in element-DOF order; compress sorts the rest;
 *  - values: element matrices scattered with SparseMatrix::add, elements in
 *    order, entry (i, j) row-major. */
or_mat *or_build(const or_space *s, int op, int acc, const or_field *coef) {
  int64_t N = s->n_dof;
  int nl = s->n_loc;
  int64_t *cnt = calloc(N, sizeof(int64_t));
  for (int64_t e = 0; e < s->n_elem; e++)
    for (int i = 0; i < nl; i++) cnt[s->elem_dof[e * nl + i]] += nl;
  int64_t maxc = 0;
  for (int64_t i = 0; i < N; i++)
    if (cnt[i] > maxc) maxc = cnt[i];
  if (maxc > N) maxc = N;
  free(cnt);
  or_sp *sp = sp_new_max(N, N, maxc);
  for (int64_t e = 0; e < s->n_elem; e++) {
    const int32_t *ed = s->elem_dof + e * nl;
    for (int i = 0; i < nl; i++)
      for (int j = 0; j < nl; j++) sp_add(sp, ed[i], ed[j]);
  }
  sp_compress(sp);
  /* shrink the column storage to the compressed size */
  sp->col = realloc(sp->col, sizeof(int64_t) * (sp->nnz > 0 ? sp->nnz : 1));
  or_mat *a = mat_new(sp);
  if (op < 0) return a; /* pattern only */
  double gv[12], K[36];
  for (int64_t e = 0; e < s->n_elem; e++) {
    elem_coords(s, e, gv);
    or_element_matrix(s->kind, op, acc, gv, coef, K);
    const int32_t *ed = s->elem_dof + e * nl;
    for (int i = 0; i < nl; i++)
      for (int j = 0; j < nl; j++) mat_add(a, ed[i], ed[j], K[i * nl + j]);
  }
  return a;
}

/* L2InnerProduct::getElementMatrix of two spaces on one element
 * (BilinearOperator.templates.h:300-326, the branch for equal elements):
 * quadrature and Jacobian of element 0, entry (j, k) accumulates
 * Jxw * phi0_j * phi1_k; K is n_loc0 x n_loc1 row-major */
void or_element_matrix_pair(int kind0, int kind1, int acc, const double *gv, double *K) {
  int tdim = kind_tdim(kind0), n0 = kind_nloc(kind0), n1 = kind_nloc(kind1);
  const qrule *q = quad_find(tdim, acc);
  double volume = tmpl_volume(tdim);
  double jac = l2g_jacobian(tdim, gv);
  for (int i = 0; i < n0 * n1; i++) K[i] = 0.0;
  for (int l = 0; l < q->n; l++) {
    double lp[3], xp[3], v0[6], v1[6];
    for (int d = 0; d < tdim; d++) lp[d] = q->d[l * (tdim + 1) + d];
    double w = q->d[l * (tdim + 1) + tdim];
    local_to_global(tdim, lp, gv, xp);
    double Jxw = w * jac * volume;
    basis_value(kind0, xp, gv, v0);
    basis_value(kind1, xp, gv, v1);
    for (int j = 0; j < n0; j++)
      for (int k = 0; k < n1; k++) K[j * n1 + k] += Jxw * v0[j] * v1[k];
  }
}

/* BilinearOperator::build over two spaces of one mesh (the same-mesh branches
 * of buildDofInfo, buildSparsityPattern and buildSparseMatrix,
 * BilinearOperator.templates.h:71-81,131-143,187-197):
 *  - coupling count of a space-0 DOF = sum over its elements of the space-1
 *    DOF count of the element; the maximum, capped at the column count, is
 *    the row length given to reinit(N0, N1, .);
 *  - each element adds (dof0_i, dof1_j) for all i, j; compress;
 *  - values: the element matrices added entry by entry, elements in order.
 * op < 0 builds the pattern only; op = OR_OP_MASS is L2InnerProduct.
 * *max_coupling (if not NULL) receives nMaxCouplingDof. */
or_mat *or_build_pair(const or_space *s0, const or_space *s1, int op, int acc, int64_t *max_coupling) {
  if (s0->n_elem != s1->n_elem) return NULL;
  int64_t N0 = s0->n_dof, N1 = s1->n_dof;
  int n0 = s0->n_loc, n1 = s1->n_loc;
  int64_t *cnt = calloc(N0 > 0 ? N0 : 1, sizeof(int64_t));
  for (int64_t e = 0; e < s0->n_elem; e++)
    for (int i = 0; i < n0; i++) cnt[s0->elem_dof[e * n0 + i]] += n1;
  int64_t maxc = 0;
  for (int64_t i = 0; i < N0; i++)
    if (cnt[i] > maxc) maxc = cnt[i];
  if (maxc > N1) maxc = N1;
  free(cnt);
  if (max_coupling) *max_coupling = maxc;
  or_sp *sp = sp_new_max(N0, N1, maxc);
  for (int64_t e = 0; e < s0->n_elem; e++) {
    const int32_t *e0 = s0->elem_dof + e * n0, *e1 = s1->elem_dof + e * n1;
    for (int i = 0; i < n0; i++)
      for (int j = 0; j < n1; j++) sp_add(sp, e0[i], e1[j]);
  }
  sp_compress(sp);
  sp->col = realloc(sp->col, sizeof(int64_t) * (sp->nnz > 0 ? sp->nnz : 1));
  or_mat *a = mat_new(sp);
  if (op < 0) return a;
  double gv[12], K[36];
  for (int64_t e = 0; e < s0->n_elem; e++) {
    elem_coords(s0, e, gv);
    or_element_matrix_pair(s0->kind, s1->kind, acc, gv, K);
    const int32_t *e0 = s0->elem_dof + e * n0, *e1 = s1->elem_dof + e * n1;
    for (int i = 0; i < n0; i++)
      for (int j = 0; j < n1; j++) mat_add(a, e0[i], e1[j], K[i * n1 + j]);
  }
  return a;
}

/* ====================================================================== */
/* right-hand side, Dirichlet elimination, error functionals               */
/* ====================================================================== */
static double u_exact(int dim, const double *x, double *grad) {
  double u = 1.0;
  for (int d = 0; d < dim; d++) u *= sin(M_PI * x[d]);
  if (grad)
    for (int d = 0; d < dim; d++) {
      double gd = M_PI * cos(M_PI * x[d]);
      for (int e = 0; e < dim; e++)
        if (e != d) gd *= sin(M_PI * x[e]);
      grad[d] = gd;
    }
  return u;
}

static double rhs_value(int fkind, int dim, const or_field *coef,
                        const double *x) {
  if (fkind == OR_F_ONE) return 1.0;
  double gu[3];
  double u = u_exact(dim, x, gu);
  if (fkind == OR_F_POISSON) return dim * M_PI * M_PI * u;
  /* -div(a grad u) = a (dim pi^2 u - grad g . grad u) for a = exp(g) */
  double gg[3];
  double g = field_g(coef, x, gg);
  double dot = 0;
  for (int d = 0; d < dim; d++) dot += gg[d] * gu[d];
  return exp(g) * (dim * M_PI * M_PI * u - dot);
}

/* Operator::L2Discretize (Operator.discretize.templates.h:121-152 for a
 * function pointer; :80-116 for a finite-element function f0, whose value at
 * a point is sum_j f0_j phi_j, FEMSpace.templates.h:1254-1265). */
void or_rhs(const or_space *s, int acc, int fkind, const or_field *coef,
            const double *nodal, double *f) {
  int tdim = s->tdim, nl = s->n_loc;
  const qrule *q = quad_find(tdim, acc);
  double volume = tmpl_volume(tdim);
  for (int64_t i = 0; i < s->n_dof; i++) f[i] = 0.0;
  for (int64_t e = 0; e < s->n_elem; e++) {
    double gv[12];
    elem_coords(s, e, gv);
    double jac = l2g_jacobian(tdim, gv);
    const int32_t *ed = s->elem_dof + e * nl;
    for (int l = 0; l < q->n; l++) {
      double lp[3], xp[3], v[6];
      for (int d = 0; d < tdim; d++) lp[d] = q->d[l * (tdim + 1) + d];
      local_to_global(tdim, lp, gv, xp);
      basis_value(s->kind, xp, gv, v);
      double fv;
      if (fkind == OR_F_NODAL) {
        fv = 0;
        for (int j = 0; j < nl; j++) fv += nodal[ed[j]] * v[j];
      } else {
        fv = rhs_value(fkind, tdim, coef, xp);
      }
      double Jxw = q->d[l * (tdim + 1) + tdim] * jac * volume;
      for (int j = 0; j < nl; j++) f[ed[j]] += Jxw * fv * v[j];
    }
  }
}

/* BoundaryConditionAdmin::apply with preserve_symmetry = true and a zero
 * Dirichlet value on every marked DOF (FEMSpace.templates.h:1338-1373):
 * f_i = A_ii * g_i; zero the off-diagonals of row i; for each column k of
 * row i, f_k -= A_ki f_i / A_ii and A_ki = 0.  u is not touched. */
void or_apply_dirichlet(const or_space *s, or_mat *a, double *u, double *f) {
  (void)u;
  const int64_t *rp = a->sp->rp, *col = a->sp->col;
  for (int64_t i = 0; i < s->n_dof; i++) {
    if (s->dof_mark[i] == 0) continue;
    f[i] = a->val[rp[i]] * 0.0;
    for (int64_t j = rp[i] + 1; j < rp[i + 1]; j++) a->val[j] = 0.0;
    for (int64_t j = rp[i] + 1; j < rp[i + 1]; j++) {
      int64_t k = col[j];
      for (int64_t p = rp[k] + 1; p < rp[k + 1]; p++)
        if (col[p] == i) {
          f[k] -= a->val[p] * f[i] / a->val[rp[i]];
          a->val[p] = 0.;
          break;
        }
    }
  }
}

/* BoundaryConditionAdmin::apply (FEMSpace.templates.h:1338-1373), preserve_
 * symmetry = true, for the DOFs whose mark is one of marks[0..n_marks) with
 * Dirichlet data g (one value per DOF): f_i = A_ii g_i, then the same
 * elimination as above.  u is not touched. */
void or_apply_dirichlet_values(const or_space *s, or_mat *a, const int32_t *marks, int n_marks,
                               const double *g, double *f) {
  or_apply_dirichlet_bc(s, a, marks, n_marks, g, f, 1);
}

/* the same with the reference's preserve_symmetry argument: without it only
 * the boundary rows change (FEMSpace.templates.h:1357-1358) */
void or_apply_dirichlet_bc(const or_space *s, or_mat *a, const int32_t *marks, int n_marks,
                           const double *g, double *f, int preserve_symmetry) {
  const int64_t *rp = a->sp->rp, *col = a->sp->col;
  for (int64_t i = 0; i < s->n_dof; i++) {
    int hit = 0;
    for (int k = 0; k < n_marks; k++) hit |= (marks[k] == s->dof_mark[i]);
    if (!hit) continue;
    f[i] = a->val[rp[i]] * g[i];
    for (int64_t j = rp[i] + 1; j < rp[i + 1]; j++) a->val[j] = 0.0;
    if (!preserve_symmetry) continue;
    for (int64_t j = rp[i] + 1; j < rp[i + 1]; j++) {
      int64_t k = col[j];
      for (int64_t p = rp[k] + 1; p < rp[k + 1]; p++)
        if (col[p] == i) {
          f[k] -= a->val[p] * f[i] / a->val[rp[i]];
          a->val[p] = 0.;
          break;
        }
    }
  }
}

/* Functional::L2Error (Functional.templates.h:296-317) and
 * Functional::H1SemiError (:399-426) against u = prod sin(pi x_d) (or 0). */
static double fe_error(const or_space *s, const double *uh, int acc, int ukind,
                       int h1) {
  int tdim = s->tdim, nl = s->n_loc;
  const qrule *q = quad_find(tdim, acc);
  double volume = tmpl_volume(tdim);
  double err = 0, nd[3] = {0, 0, 0};
  for (int64_t e = 0; e < s->n_elem; e++) {
    double gv[12];
    elem_coords(s, e, gv);
    double jac = l2g_jacobian(tdim, gv);
    const int32_t *ed = s->elem_dof + e * nl;
    for (int l = 0; l < q->n; l++) {
      double lp[3], xp[3], v[6], g[18], ge[3] = {0, 0, 0};
      for (int d = 0; d < tdim; d++) lp[d] = q->d[l * (tdim + 1) + d];
      local_to_global(tdim, lp, gv, xp);
      double Jxw = q->d[l * (tdim + 1) + tdim] * jac * volume;
      double ue = ukind == OR_U_ZERO ? 0.0 : u_exact(tdim, xp, ge);
      if (!h1) {
        basis_value(s->kind, xp, gv, v);
        double val = 0;
        for (int j = 0; j < nl; j++) val += uh[ed[j]] * v[j];
        double df = ue - val;
        err += Jxw * df * df;
      } else {
        basis_grad(s->kind, xp, gv, g);
        double gh[3] = {0, 0, 0};
        for (int j = 0; j < nl; j++)
          for (int d = 0; d < tdim; d++) gh[d] += uh[ed[j]] * g[j * tdim + d];
        if (ukind == OR_U_ZERO) ge[0] = ge[1] = ge[2] = 0;
        for (int d = 0; d < tdim; d++) {
          double df = gh[d] - ge[d];
          nd[d] += Jxw * df * df;
        }
      }
    }
  }
  if (h1) {
    err = 0;
    for (int d = 0; d < tdim; d++) err += nd[d];
  }
  return sqrt(fabs(err));
}

double or_l2_error(const or_space *s, const double *uh, int acc, int ukind) {
  return fe_error(s, uh, acc, ukind, 0);
}
double or_h1_error(const or_space *s, const double *uh, int acc, int ukind) {
  return fe_error(s, uh, acc, ukind, 1);
}

/* ====================================================================== */
/* AMG                                                                     */
/* ====================================================================== */
#define OR_MAX_LEVELS 64
struct or_amg {
  int lazy;
  int n_project;
  int full;
  double alpha;
  unsigned nmpo;
  double nmpsd;
  const or_mat *A[OR_MAX_LEVELS + 1]; /* A[0] borrowed */
  or_mat *P[OR_MAX_LEVELS];
  or_mat *Pt[OR_MAX_LEVELS];
  int64_t nc;
  double *minv;
};

/* AMGSolver::getPMPt (src/AMGSolver.cpp:288-358): row i of P A P^T collects
 * its columns in discovery order over (P row i) x (A row a) x (P^T row b)
 * with the diagonal first; the pattern is then compressed (diagonal first,
 * rest ascending) and the values are the accumulated triple products. */
static or_mat *get_pmpt(const or_mat *P, const or_mat *M, const or_mat *Pt) {
  int64_t nc = P->sp->m;
  const int64_t *Prp = P->sp->rp, *Pc = P->sp->col;
  const int64_t *Mrp = M->sp->rp, *Mc = M->sp->col;
  const int64_t *Trp = Pt->sp->rp, *Tc = Pt->sp->col;
  char *flag = malloc(nc > 0 ? nc : 1);
  memset(flag, 1, nc);
  int64_t *index = malloc(sizeof(int64_t) * (nc > 0 ? nc : 1));
  double *row_entry = calloc(nc > 0 ? nc : 1, sizeof(double));
  int64_t cap = nc * 8 + 16, used = 0;
  int64_t *col = malloc(sizeof(int64_t) * cap);
  double *val = malloc(sizeof(double) * cap);
  int64_t *rp = malloc(sizeof(int64_t) * (nc + 1));
  rp[0] = 0;
  for (int64_t i = 0; i < nc; i++) {
    int64_t len = 1;
    flag[i] = 0;
    index[0] = i;
    for (int64_t j = Prp[i]; j < Prp[i + 1]; j++) {
      int64_t a = Pc[j];
      for (int64_t k = Mrp[a]; k < Mrp[a + 1]; k++) {
        int64_t b = Mc[k];
        for (int64_t l = Trp[b]; l < Trp[b + 1]; l++) {
          int64_t c = Tc[l];
          if (flag[c]) {
            index[len++] = c;
            flag[c] = 0;
          }
        }
      }
    }
    for (int64_t j = Prp[i]; j < Prp[i + 1]; j++) {
      int64_t a = Pc[j];
      for (int64_t k = Mrp[a]; k < Mrp[a + 1]; k++) {
        int64_t b = Mc[k];
        for (int64_t l = Trp[b]; l < Trp[b + 1]; l++)
          row_entry[Tc[l]] += P->val[j] * M->val[k] * Pt->val[l];
      }
    }
    if (used + len > cap) {
      while (used + len > cap) cap *= 2;
      col = realloc(col, sizeof(int64_t) * cap);
      val = realloc(val, sizeof(double) * cap);
    }
    /* compress: diagonal first, the rest ascending */
    col[used] = i;
    for (int64_t j = 1; j < len; j++) col[used + j] = index[j];
    qsort(col + used + 1, len - 1, sizeof(int64_t), i64_cmp);
    for (int64_t j = 0; j < len; j++) {
      val[used + j] = row_entry[col[used + j]];
      row_entry[col[used + j]] = 0.0;
      flag[col[used + j]] = 1;
    }
    used += len;
    rp[i + 1] = used;
  }
  or_mat *A = or_mat_from_csr(nc, nc, rp, col, val);
  free(flag);
  free(index);
  free(row_entry);
  free(col);
  free(val);
  free(rp);
  return A;
}

/* AMGSolver::Project (src/AMGSolver.cpp:18-226).
 *  1. C/F split, greedy in index order: beta_i = alpha * max_{j!=i}|a_ij|;
 *     a candidate row becomes C and marks its strong (|a_ij| >= beta_i)
 *     non-C neighbours F (:38-64).
 *  2. an F point with a strong F neighbour k > i that shares no C point with
 *     the strong C points of i becomes C (:71-122).
 *  3. P (coarse x fine) and P^T patterns: each C point, plus every non-C
 *     neighbour of a C point (:127-163) -- weak and F-F couplings keep their
 *     slots, so P carries explicit structural zeros.
 *  4. values: C rows 1; F rows |a_il / a~_ii| with a~_ii = a_ii + sum of weak
 *     a_ij for strong C neighbours l, and the distance-two formula through a
 *     strong F neighbour l over the common C points (:174-219).
 *  5. A_c = getPMPt(P, A, P^T). */
static void amg_project(const or_mat *M, double alpha, or_mat **Pout,
                        or_mat **PMPt, or_mat **Ptout) {
  int64_t n = M->sp->m;
  const int64_t *rp = M->sp->rp, *col = M->sp->col;
  const double *v = M->val;
  int *flag = calloc(n, sizeof(int));
  double *beta = calloc(n, sizeof(double));
  int64_t n_core = 0;
  for (int64_t i = 0; i < n; i++) {
    beta[i] = 0.0;
    for (int64_t j = rp[i] + 1; j < rp[i + 1]; j++)
      beta[i] = fabs(v[j]) < beta[i] ? beta[i] : fabs(v[j]);
    beta[i] *= alpha;
    if (flag[i] != 0) continue;
    for (int64_t j = rp[i] + 1; j < rp[i + 1]; j++) {
      int64_t k = col[j];
      if (flag[k] == -1) continue;
      if (fabs(v[j]) < beta[i]) continue;
      flag[k] = 1;
    }
    flag[i] = -1;
    n_core++;
  }
  /* dep_pnt membership through a stamped marker (the reference scans a list) */
  int64_t *mark = malloc(sizeof(int64_t) * n);
  for (int64_t i = 0; i < n; i++) mark[i] = -1;
  for (int64_t i = 0; i < n; i++) {
    if (flag[i] != 1) continue;
    for (int64_t j = rp[i] + 1; j < rp[i + 1]; j++) {
      int64_t k = col[j];
      if (fabs(v[j]) < beta[i]) continue;
      if (flag[k] == -1) mark[k] = i;
    }
    int is_FF = 0;
    for (int64_t j = rp[i] + 1; j < rp[i + 1]; j++) {
      int64_t k = col[j];
      if (flag[k] == -1 || k < i || fabs(v[j]) < beta[i]) continue;
      is_FF = 1;
      for (int64_t j1 = rp[k] + 1; j1 < rp[k + 1]; j1++) {
        int64_t k1 = col[j1];
        if (flag[k1] != -1) continue;
        if (mark[k1] == i) {
          is_FF = 0;
          break;
        }
      }
      if (is_FF) break;
    }
    if (!is_FF) continue;
    flag[i] = -1;
    n_core++;
  }
  int64_t *core_point = malloc(sizeof(int64_t) * (n_core > 0 ? n_core : 1));
  int64_t *core_at = malloc(sizeof(int64_t) * n);
  int64_t *row_length = calloc(n_core > 0 ? n_core : 1, sizeof(int64_t));
  int64_t *col_length = calloc(n, sizeof(int64_t));
  for (int64_t i = 0, c = 0; i < n; i++) {
    core_at[i] = -1;
    if (flag[i] != -1) continue;
    core_point[c] = i;
    core_at[i] = c;
    for (int64_t j = rp[i]; j < rp[i + 1]; j++) {
      row_length[c]++;
      col_length[col[j]]++;
    }
    c++;
  }
  or_sp *spP = sp_new_lengths(n_core, n, row_length);
  or_sp *spPt = sp_new_lengths(n, n_core, col_length);
  for (int64_t i = 0; i < n_core; i++) {
    int64_t k = core_point[i];
    sp_add(spP, i, k);
    sp_add(spPt, k, i);
    for (int64_t j = rp[k] + 1; j < rp[k + 1]; j++) {
      int64_t l = col[j];
      if (flag[l] == -1) continue;
      sp_add(spP, i, l);
      sp_add(spPt, l, i);
    }
  }
  sp_compress(spP);
  sp_compress(spPt);
  or_mat *P = mat_new(spP);
  or_mat *Pt = mat_new(spPt);
  const int64_t *Trp = spPt->rp, *Tc = spPt->col;
  int64_t *dep = malloc(sizeof(int64_t) * (n_core > 0 ? n_core : 1));
  for (int64_t c = 0; c < n_core; c++) dep[c] = -1;
  for (int64_t i = 0; i < n; i++) {
    int64_t k = core_at[i];
    if (k >= 0) {
      mat_add(Pt, i, k, 1.0);
      mat_add(P, k, i, 1.0);
      continue;
    }
    double aii = v[rp[i]];
    for (int64_t j = rp[i] + 1; j < rp[i + 1]; j++)
      if (fabs(v[j]) < beta[i]) aii += v[j];
    for (int64_t i1 = Trp[i]; i1 < Trp[i + 1]; i1++) dep[Tc[i1]] = i;
    for (int64_t j = rp[i] + 1; j < rp[i + 1]; j++) {
      int64_t l = col[j];
      if (fabs(v[j]) < beta[i]) continue;
      if (core_at[l] >= 0) {
        double w = fabs(v[j] / aii);
        mat_add(Pt, i, core_at[l], w);
        mat_add(P, core_at[l], i, w);
      } else {
        double d = 0.0;
        for (int64_t i1 = Trp[l]; i1 < Trp[l + 1]; i1++) {
          int64_t c1 = Tc[i1];
          if (dep[c1] == i) d += mat_el(M, l, core_point[c1]);
        }
        if (fabs(d * aii) < 1e-16) continue;
        for (int64_t i1 = Trp[l]; i1 < Trp[l + 1]; i1++) {
          int64_t c1 = Tc[i1];
          if (dep[c1] == i) {
            double w = fabs(mat_el(M, i, l) * mat_el(M, l, core_point[c1]) / (d * aii));
            mat_add(Pt, i, c1, w);
            mat_add(P, c1, i, w);
          }
        }
      }
    }
  }
  *Pout = P;
  *PMPt = get_pmpt(P, M, Pt);
  *Ptout = Pt; /* kept for inspection; the reference discards it here */
  free(dep);
  free(flag);
  free(beta);
  free(mark);
  free(core_point);
  free(core_at);
  free(row_length);
  free(col_length);
}

/* AMGSolver::lazyProject (src/AMGSolver.cpp:235-284): a row becomes a core
 * point unless an earlier core row already contains it; P has the core point
 * (weight 1) and each neighbour l with weight 1/counter[l]. */
static void amg_lazy_project(const or_mat *M, or_mat **Pout, or_mat **PMPt,
                             or_mat **Ptout) {
  int64_t n = M->sp->m;
  const int64_t *rp = M->sp->rp, *col = M->sp->col;
  uint32_t *counter = calloc(n, sizeof(uint32_t));
  int64_t *core_point = malloc(sizeof(int64_t) * n);
  int64_t n_core = 0;
  for (int64_t i = 0; i < n; i++) {
    if (counter[i] > 0) continue;
    for (int64_t j = rp[i] + 1; j < rp[i + 1]; j++) counter[col[j]]++;
    core_point[n_core++] = i;
  }
  int64_t maxlen = sp_max_entries_per_row(M->sp);
  or_sp *spP = sp_new_max(n_core, M->sp->n, maxlen);
  or_sp *spPt = sp_new_max(M->sp->n, n_core, maxlen);
  for (int64_t i = 0; i < n_core; i++) {
    int64_t k = core_point[i];
    sp_add(spP, i, k);
    sp_add(spPt, k, i);
    for (int64_t j = rp[k] + 1; j < rp[k + 1]; j++) {
      sp_add(spP, i, col[j]);
      sp_add(spPt, col[j], i);
    }
  }
  sp_compress(spP);
  sp_compress(spPt);
  or_mat *P = mat_new(spP), *Pt = mat_new(spPt);
  for (int64_t i = 0; i < n_core; i++) {
    int64_t k = core_point[i];
    mat_add(P, i, k, 1.0);
    mat_add(Pt, k, i, 1.0);
    for (int64_t j = rp[k] + 1; j < rp[k + 1]; j++) {
      int64_t l = col[j];
      mat_add(P, i, l, 1.0 / counter[l]);
      mat_add(Pt, l, i, 1.0 / counter[l]);
    }
  }
  *Pout = P;
  *Ptout = Pt;
  *PMPt = get_pmpt(P, M, Pt);
  free(counter);
  free(core_point);
}

/* Gauss-Jordan inverse with partial pivoting of a row-major n x n matrix.
 * Stands in for DenseMatrix::gauss_jordan = LAPACKE dgetrf + dgetri
 * (include/AFEPack/DenseMatrix.h:131-172); returns -1 when singular. */
int or_dense_inverse(int n, double *a) {
  int *perm = malloc(sizeof(int) * n);
  for (int i = 0; i < n; i++) perm[i] = i;
  double *b = calloc((size_t)n * n, sizeof(double));
  for (int i = 0; i < n; i++) b[i * n + i] = 1.0;
  for (int c = 0; c < n; c++) {
    int p = c;
    for (int r = c + 1; r < n; r++)
      if (fabs(a[r * n + c]) > fabs(a[p * n + c])) p = r;
    if (a[p * n + c] == 0.0) {
      free(perm);
      free(b);
      return -1;
    }
    if (p != c)
      for (int j = 0; j < n; j++) {
        double t = a[c * n + j];
        a[c * n + j] = a[p * n + j];
        a[p * n + j] = t;
        t = b[c * n + j];
        b[c * n + j] = b[p * n + j];
        b[p * n + j] = t;
      }
    double piv = a[c * n + c];
    for (int j = 0; j < n; j++) {
      a[c * n + j] /= piv;
      b[c * n + j] /= piv;
    }
    for (int r = 0; r < n; r++) {
      if (r == c) continue;
      double f = a[r * n + c];
      if (f == 0.0) continue;
      for (int j = 0; j < n; j++) {
        a[r * n + j] -= f * a[c * n + j];
        b[r * n + j] -= f * b[c * n + j];
      }
    }
  }
  memcpy(a, b, sizeof(double) * n * n);
  free(perm);
  free(b);
  return 0;
}

static void amg_coarse_inverse(or_amg *g) {
  const or_mat *Ac = g->A[g->n_project];
  int64_t n = Ac->sp->m;
  free(g->minv);
  g->nc = n;
  g->minv = calloc((size_t)(n * n > 0 ? n * n : 1), sizeof(double));
  for (int64_t k = 0; k < n; k++)
    for (int64_t m = Ac->sp->rp[k]; m < Ac->sp->rp[k + 1]; m++)
      g->minv[k * n + Ac->sp->col[m]] = Ac->val[m];
  or_dense_inverse((int)n, g->minv);
}

/* AMGSolver::reinit (src/AMGSolver.cpp:384-438) and lazyInit (:442-496):
 * project until the order is <= nmpo and the density nnz/m^2 (in float) is
 * >= nmpsd, or until a projection no longer reduces the order (then the
 * coarsest level is solved by diagonal scaling); otherwise the coarsest
 * matrix is inverted densely. */
or_amg *or_amg_new(const or_mat *M, int lazy, double alpha, unsigned nmpo,
                   double nmpsd) {
  or_amg *g = calloc(1, sizeof *g);
  g->lazy = lazy;
  g->alpha = alpha;
  g->nmpo = nmpo;
  g->nmpsd = nmpsd;
  uint32_t k, m;
  k = m = (uint32_t)M->sp->m;
  float sparse_degree = (float)M->sp->nnz / ((float)m * (float)m);
  g->A[0] = M;
  g->n_project = 0;
  g->full = 1;
  while (m > nmpo || sparse_degree < nmpsd) {
    or_mat *P, *PMPt, *Pt = NULL;
    if (lazy)
      amg_lazy_project(g->A[g->n_project], &P, &PMPt, &Pt);
    else
      amg_project(g->A[g->n_project], alpha, &P, &PMPt, &Pt);
    m = (uint32_t)PMPt->sp->m;
    if (k == m || g->n_project >= OR_MAX_LEVELS) {
      g->full = 0;
      or_mat_free(P);
      or_mat_free(PMPt);
      or_mat_free(Pt);
      break;
    }
    k = m;
    g->P[g->n_project] = P;
    g->Pt[g->n_project] = Pt;
    g->A[g->n_project + 1] = PMPt;
    sparse_degree = (float)PMPt->sp->nnz / ((float)m * (float)m);
    g->n_project++;
  }
  if (g->full) amg_coarse_inverse(g);
  return g;
}

/* AMGSolver::lazyReinit (src/AMGSolver.cpp:505-545): same pattern object
 * required (returns -1 where the reference aborts); P and P^T are kept and
 * only the Galerkin products and the coarse inverse are recomputed. */
int or_amg_lazy_reinit(or_amg *g, const or_mat *M) {
  if (g->A[0]->sp != M->sp) return -1;
  g->A[0] = M;
  for (int j = 0; j < g->n_project; j++) {
    or_mat_free((or_mat *)g->A[j + 1]);
    g->A[j + 1] = get_pmpt(g->P[j], g->A[j], g->Pt[j]);
  }
  if (g->full) amg_coarse_inverse(g);
  return 0;
}

void or_amg_free(or_amg *g) {
  if (!g) return;
  for (int i = 0; i < g->n_project; i++) {
    or_mat_free(g->P[i]);
    or_mat_free(g->Pt[i]);
    or_mat_free((or_mat *)g->A[i + 1]);
  }
  free(g->minv);
  free(g);
}

int or_amg_n_project(const or_amg *g) { return g->n_project; }
int or_amg_coarse_full(const or_amg *g) { return g->full; }
const or_mat *or_amg_A(const or_amg *g, int l) { return g->A[l]; }
const or_mat *or_amg_P(const or_amg *g, int l) { return g->P[l]; }
const or_mat *or_amg_Pt(const or_amg *g, int l) { return g->Pt[l]; }
void or_amg_coarse_inverse(const or_amg *g, double *minv) {
  if (g->minv) memcpy(minv, g->minv, sizeof(double) * g->nc * g->nc);
}

/* AMGSolver::GaussSidel (src/AMGSolver.cpp:362-381): s forward lexicographic
 * sweeps in place; the diagonal is the first entry of each row. */
void or_gauss_seidel(const or_mat *M, double *x, const double *r, int s) {
  const int64_t *rp = M->sp->rp, *col = M->sp->col;
  for (int i = 0; i < s; i++)
    for (int64_t j = 0; j < M->sp->m; j++) {
      double r0 = r[j];
      for (int64_t k = rp[j] + 1; k < rp[j + 1]; k++) r0 -= x[col[k]] * M->val[k];
      x[j] = r0 / M->val[rp[j]];
    }
}

static double l1_norm(const double *v, int64_t n) {
  double s = 0;
  for (int64_t i = 0; i < n; i++) s += fabs(v[i]);
  return s;
}

/* AMGSolver::solve (src/AMGSolver.cpp:549-639): exactly step+1 V-cycles
 * from the initial guess x (no early exit).  Down: GS(s), residual, restrict
 * with P, zero the coarse guess.  Coarsest: dense inverse (or diagonal
 * scaling when coarsening stalled; GS when not solving exactly).  Up:
 * x_l += P_l^T x_{l+1}, GS(s).  res_hist[0] and res_hist[c+1] are the L1
 * residuals ||A x - r||_1 before and after cycle c. */
void or_amg_solve(const or_amg *g, double *x, const double *r, unsigned smooth,
                  int64_t step, int exactly, double *res_hist) {
  int np = g->n_project;
  int64_t n0 = g->A[0]->sp->m;
  double *px[OR_MAX_LEVELS + 1], *pr[OR_MAX_LEVELS + 1];
  px[0] = x;
  pr[0] = malloc(sizeof(double) * n0);
  memcpy(pr[0], r, sizeof(double) * n0);
  for (int i = 1; i <= np; i++) {
    px[i] = calloc(g->A[i]->sp->m, sizeof(double));
    pr[i] = calloc(g->A[i]->sp->m, sizeof(double));
  }
  double *v0 = malloc(sizeof(double) * n0), *v1 = malloc(sizeof(double) * n0);
  double *v2 = malloc(sizeof(double) * n0);
  or_mat_vmult(g->A[0], x, v2);
  for (int64_t i = 0; i < n0; i++) v2[i] += -1.0 * r[i];
  if (res_hist) res_hist[0] = l1_norm(v2, n0);
  int64_t it = 0;
  while (1) {
    for (int i = 0; i < np; i++) {
      int64_t ni = g->A[i]->sp->m;
      or_gauss_seidel(g->A[i], px[i], pr[i], (int)smooth);
      or_mat_vmult(g->A[i], px[i], v0);
      for (int64_t j = 0; j < ni; j++) v1[j] = pr[i][j] + -1.0 * v0[j];
      or_mat_vmult(g->P[i], v1, pr[i + 1]);
      memset(px[i + 1], 0, sizeof(double) * g->A[i + 1]->sp->m);
    }
    const or_mat *Ac = g->A[np];
    int64_t nc = Ac->sp->m;
    if (exactly) {
      if (g->full) {
        for (int64_t a = 0; a < nc; a++) {
          double s = 0;
          for (int64_t b = 0; b < nc; b++) s += g->minv[a * nc + b] * pr[np][b];
          px[np][a] = s;
        }
      } else {
        for (int64_t j = 0; j < nc; j++) px[np][j] = pr[np][j] / Ac->val[Ac->sp->rp[j]];
      }
    } else {
      or_gauss_seidel(Ac, px[np], pr[np], (int)smooth);
    }
    for (int i = np - 1; i >= 0; i--) {
      int64_t ni = g->A[i]->sp->m;
      or_mat_tvmult(g->P[i], px[i + 1], v0);
      for (int64_t j = 0; j < ni; j++) px[i][j] += 1.0 * v0[j];
      or_gauss_seidel(g->A[i], px[i], pr[i], (int)smooth);
    }
    or_mat_vmult(g->A[0], x, v2);
    for (int64_t i = 0; i < n0; i++) v2[i] += -1.0 * r[i];
    it++;
    if (res_hist) res_hist[it] = l1_norm(v2, n0);
    if (it > step) break;
  }
  for (int i = 0; i <= np; i++) {
    if (i > 0) free(px[i]);
    free(pr[i]);
  }
  free(v0);
  free(v1);
  free(v2);
}
