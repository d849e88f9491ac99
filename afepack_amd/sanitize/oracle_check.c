/* Drives the oracle (oracle/afe_oracle.c) over small meshes for a build with
 * AddressSanitizer and UndefinedBehaviorSanitizer: spaces with and without
 * side lists, every operator, the two-space build, right-hand sides,
 * Dirichlet elimination, both coarsenings, lazy reinit, the stationary solve
 * and the error functionals.  Exits 0 when the solves converge. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "../../oracle/afe_oracle.h"

/* n x n squares split along the diagonal, vertices marked 1 on the boundary */
static void square(int n, double *c, int32_t *ev, int32_t *vm) {
  for (int j = 0; j <= n; j++)
    for (int i = 0; i <= n; i++) {
      int v = j * (n + 1) + i;
      c[2 * v] = (double)i / n + (i > 0 && i < n ? 0.1 / n * sin(7.0 * v) : 0.0);
      c[2 * v + 1] = (double)j / n + (j > 0 && j < n ? 0.1 / n * cos(5.0 * v) : 0.0);
      vm[v] = (i == 0 || j == 0 || i == n || j == n);
    }
  int e = 0;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) {
      int a = j * (n + 1) + i, b = a + 1, d = a + n + 1, f = d + 1;
      ev[3 * e] = a, ev[3 * e + 1] = b, ev[3 * e + 2] = f, e++;
      ev[3 * e] = a, ev[3 * e + 1] = f, ev[3 * e + 2] = d, e++;
    }
}

/* n^3 cubes, each split into the 6 tetrahedra around its main diagonal,
 * every tetrahedron positively oriented */
static void cube(int n, double *c, int32_t *ev, int32_t *vm) {
  int m = n + 1;
  for (int k = 0; k <= n; k++)
    for (int j = 0; j <= n; j++)
      for (int i = 0; i <= n; i++) {
        int v = (k * m + j) * m + i;
        c[3 * v] = (double)i / n, c[3 * v + 1] = (double)j / n, c[3 * v + 2] = (double)k / n;
        vm[v] = (i == 0 || j == 0 || k == 0 || i == n || j == n || k == n);
      }
  static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  int e = 0;
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++)
        for (int t = 0; t < 6; t++) {
          int p[3] = {i, j, k}, q[4];
          q[0] = (k * m + j) * m + i;
          for (int s = 0; s < 3; s++) {
            p[perm[t][s]]++;
            q[s + 1] = (p[2] * m + p[1]) * m + p[0];
          }
          const double *x0 = c + 3 * q[0], *x1 = c + 3 * q[1], *x2 = c + 3 * q[2], *x3 = c + 3 * q[3];
          double a[3], b[3], d[3];
          for (int s = 0; s < 3; s++) a[s] = x1[s] - x0[s], b[s] = x2[s] - x0[s], d[s] = x3[s] - x0[s];
          double det = a[0] * (b[1] * d[2] - b[2] * d[1]) - a[1] * (b[0] * d[2] - b[2] * d[0]) +
                       a[2] * (b[0] * d[1] - b[1] * d[0]);
          if (det < 0) {
            int tmp = q[2];
            q[2] = q[3];
            q[3] = tmp;
          }
          for (int s = 0; s < 4; s++) ev[4 * e + s] = q[s];
          e++;
        }
}

static int solve_and_check(const or_space *s, int kind, int lazy) {
  int acc = kind == OR_P2_TRI ? 6 : 4;
  int64_t N = or_space_n_dof(s);
  or_mat *A = or_build(s, OR_OP_STIFF, acc, NULL);
  double *f = calloc(N, sizeof(double)), *u = calloc(N, sizeof(double)), *x = calloc(N, sizeof(double));
  double hist[64];
  or_rhs(s, acc, OR_F_POISSON, NULL, NULL, f);
  or_apply_dirichlet(s, A, u, f);
  or_amg *g = or_amg_new(A, lazy, 0.25, 20, 0.382);
  if (lazy && or_amg_lazy_reinit(g, A) != 0) return 1;
  or_amg_solve(g, x, f, 3, 40, 1, hist);
  double e2 = or_l2_error(s, x, acc, OR_U_SINPROD), e1 = or_h1_error(s, x, acc, OR_U_SINPROD);
  int bad = !(e2 < 0.1 && e1 < 2.0 && e2 > 0.0);
  if (bad) fprintf(stderr, "kind %d lazy %d: L2 %g H1 %g\n", kind, lazy, e2, e1);
  or_amg_free(g);
  or_mat_free(A);
  free(f), free(u), free(x);
  return bad;
}

int main(void) {
  int fail = 0;
  const int n2 = 12, n3 = 4;
  double c2[(n2 + 1) * (n2 + 1) * 2], c3[(n3 + 1) * (n3 + 1) * (n3 + 1) * 3];
  int32_t ev2[2 * n2 * n2 * 3], vm2[(n2 + 1) * (n2 + 1)], ev3[6 * n3 * n3 * n3 * 4],
      vm3[(n3 + 1) * (n3 + 1) * (n3 + 1)];
  square(n2, c2, ev2, vm2);
  cube(n3, c3, ev3, vm3);
  /* every edge as a side (the EasyMesh .s list holds them all): interior
   * edges marked 0, boundary edges 2 + the side of the square they lie on */
  enum { NS = 3 * n2 * n2 + 2 * n2 };
  int32_t sv[2 * NS], sm[NS];
  int ns = 0;
  for (int j = 0; j <= n2; j++)
    for (int i = 0; i <= n2; i++) {
      int v = j * (n2 + 1) + i;
      if (i < n2) sv[2 * ns] = v, sv[2 * ns + 1] = v + 1, sm[ns++] = j == 0 ? 2 : (j == n2 ? 3 : 0);
      if (j < n2) sv[2 * ns] = v, sv[2 * ns + 1] = v + n2 + 1, sm[ns++] = i == 0 ? 4 : (i == n2 ? 5 : 0);
      if (i < n2 && j < n2) sv[2 * ns] = v, sv[2 * ns + 1] = v + n2 + 2, sm[ns++] = 0;
    }
  or_space *p1 = or_space_new(OR_P1_TRI, (n2 + 1) * (n2 + 1), c2, 2 * n2 * n2, ev2, vm2);
  or_space *p2 = or_space_new_sides(OR_P2_TRI, (n2 + 1) * (n2 + 1), c2, 2 * n2 * n2, ev2, vm2, ns, sv, sm);
  or_space *t1 = or_space_new(OR_P1_TET, (n3 + 1) * (n3 + 1) * (n3 + 1), c3, 6 * n3 * n3 * n3, ev3, vm3);
  if (!p1 || !p2 || !t1 || ns != NS) {
    fprintf(stderr, "space construction failed\n");
    return 1;
  }
  for (int lazy = 0; lazy <= 1; lazy++) {
    fail |= solve_and_check(p1, OR_P1_TRI, lazy);
    fail |= solve_and_check(p2, OR_P2_TRI, lazy);
    fail |= solve_and_check(t1, OR_P1_TET, lazy);
  }
  double k[8] = {1, 2, 3, 1, 2, 1, 1, 3}, phi[4] = {0.1, 0.2, 0.3, 0.4}, w[4] = {0.5, 0.25, 0.2, 0.1};
  or_field fld = {4, 2, k, phi, w, 0.3};
  or_mat *V = or_build(p2, OR_OP_VARSTIFF, 6, &fld), *M = or_build(p1, OR_OP_MASS, 4, NULL);
  int64_t mc = 0;
  or_mat *X = or_build_pair(p1, p2, OR_OP_MASS, 4, &mc), *Y = or_build_pair(p2, p1, OR_OP_MASS, 4, NULL);
  int64_t n1 = or_space_n_dof(p1), nn2 = or_space_n_dof(p2);
  double *a = calloc(nn2, sizeof(double)), *b = calloc(n1, sizeof(double)), *cc = calloc(nn2, sizeof(double));
  for (int64_t i = 0; i < nn2; i++) a[i] = 1.0;
  or_mat_vmult(X, a, b);   /* row sums: int phi0_i */
  or_mat_tvmult(X, b, cc); /* and back */
  double area = 0;
  for (int64_t i = 0; i < n1; i++) area += b[i];
  if (fabs(area - 1.0) > 1e-12 || or_mat_m(X) != n1 || or_mat_n(X) != nn2 || mc <= 0) {
    fprintf(stderr, "two-space mass: area %.17g\n", area);
    fail = 1;
  }
  or_mat_free(V), or_mat_free(M), or_mat_free(X), or_mat_free(Y);
  free(a), free(b), free(cc);
  or_space_free(p1), or_space_free(p2), or_space_free(t1);
  printf(fail ? "oracle_check: FAILED\n" : "oracle_check: ok\n");
  return fail;
}
