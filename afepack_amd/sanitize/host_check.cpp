// Drives the host side of the library (csrc/host.cpp) over small meshes of
// every element kind for a build with AddressSanitizer and
// UndefinedBehaviorSanitizer: mesh generation, first-touch numbering with and
// without side lists, the two-space pattern, both coarsenings (Project and
// lazyProject), lazy refresh, colouring and the sliced-ELL layouts with
// their source maps.  Exits 0 when every structural check holds.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <vector>

#include "../csrc/host.hpp"

using namespace afe;

static int g_fail = 0;
#define CHECK(c)                                                        \
  do {                                                                  \
    if (!(c)) {                                                         \
      std::fprintf(stderr, "check failed at %s:%d: %s\n", __FILE__, __LINE__, #c); \
      g_fail = 1;                                                       \
    }                                                                   \
  } while (0)

// a diagonally dominant matrix on the pattern (stiffness-like signs)
static Csr laplacian_on(const Space &s) {
  Csr A;
  A.m = A.n = s.n_dof;
  A.rp = s.rp;
  A.col = s.col;
  A.val.assign(A.rp[A.m], 0.0);
  for (int64_t i = 0; i < A.m; i++) {
    const int64_t b = A.rp[i], e = A.rp[i + 1];
    for (int64_t k = b + 1; k < e; k++) A.val[k] = -1.0 - 0.01 * (double)((i + A.col[k]) % 7);
    double d = 0.1;
    for (int64_t k = b + 1; k < e; k++) d -= A.val[k];
    A.val[b] = d;
  }
  return A;
}

static void layouts(const Hierarchy &h) {
  for (size_t l = 0; l < h.A.size(); l++) {
    const Csr &A = h.A[l];
    std::vector<int32_t> colour;
    const int nc = colour_rows(A, colour, l == 0 ? DSATUR_MAX_ROWS : DSATUR_MAX_ROWS_COARSE);
    for (int64_t i = 0; i < A.m; i++)
      for (int64_t k = A.rp[i]; k < A.rp[i + 1]; k++)
        if (A.col[k] != i) CHECK(colour[i] != colour[A.col[k]]);
    std::vector<int32_t> perm(A.m), iperm(A.m), groups(nc, 0);
    for (int64_t i = 0; i < A.m; i++) groups[colour[i]]++;
    std::vector<int64_t> start(nc + 1, 0);
    for (int c = 0; c < nc; c++) start[c + 1] = start[c] + groups[c];
    for (int64_t i = 0; i < A.m; i++) perm[start[colour[i]]++] = (int32_t)i;
    for (int64_t i = 0; i < A.m; i++) iperm[perm[i]] = (int32_t)i;
    for (int gs = 0; gs <= 2; gs++) {
      Sell out;
      std::vector<int32_t> smap, dsrc;
      build_sell(A, perm, iperm, groups, true, out, &smap, &dsrc, gs);
      CHECK((int64_t)smap.size() == A.nnz());
      CHECK((int64_t)dsrc.size() == A.m);
      for (int64_t k = 0; k < A.nnz(); k++) CHECK(smap[k] < (int32_t)out.cols.size());
    }
    if (l + 1 < h.A.size()) {
      Sell sp, st;
      const int64_t nc1 = h.A[l + 1].m;
      std::vector<int32_t> cperm(nc1);
      for (int64_t i = 0; i < nc1; i++) cperm[i] = (int32_t)i;
      build_sell(h.P[l], cperm, iperm, {(int32_t)h.P[l].m}, false, sp, nullptr, nullptr,
                 row_lane_shift(h.P[l].m, 1, h.P[l].nnz(), MAX_ROW_LANE_SHIFT));
      build_sell(transpose(h.P[l]), perm, cperm, {(int32_t)A.m}, false, st);
      CHECK(sp.n_rows == nc1 && st.n_rows == A.m);
    }
  }
}

int main() {
  set_host_threads(2);
  for (int kind : {P1_TRI, P2_TRI, P1_TET}) {
    const int n = kind == P1_TET ? 4 : 10;
    const int64_t nv = kind == P1_TET ? (int64_t)(n + 1) * (n + 1) * (n + 1) : (int64_t)(n + 1) * (n + 1);
    const int64_t ne = kind == P1_TET ? 6LL * n * n * n : 2LL * n * n;
    const int nvtx = kind_nvtx(kind), dim = kind_tdim(kind);
    std::vector<double> coords(nv * dim);
    std::vector<int32_t> ev(ne * nvtx), vmark(nv);
    if (kind == P1_TET)
      mesh_cube(n, 0.05, 3, coords.data(), ev.data(), vmark.data());
    else
      mesh_square(n, 0.2, 3, coords.data(), ev.data(), vmark.data());
    Space s;
    space_build(s, kind, nv, coords.data(), ne, ev.data(), vmark.data());
    CHECK(s.n_dof > 0 && (int64_t)s.rp.size() == s.n_dof + 1);
    for (int64_t i = 0; i < s.n_dof; i++) CHECK(s.col[s.rp[i]] == i);
    if (dim == 2) {
      // the boundary edges as a side list, marked by their lower vertex
      std::map<std::pair<int, int>, int> cnt;
      for (int64_t e = 0; e < ne; e++)
        for (int k = 0; k < 3; k++) {
          int a = ev[e * 3 + (k + 1) % 3], b = ev[e * 3 + (k + 2) % 3];
          cnt[{std::min(a, b), std::max(a, b)}]++;
        }
      std::vector<int32_t> sv, sm;
      for (auto &kv : cnt) {
        sv.push_back(kv.first.first);
        sv.push_back(kv.first.second);
        sm.push_back(kv.second == 1 ? 1 + kv.first.first % 4 : 0);
      }
      Space s2;
      space_build(s2, kind, nv, coords.data(), ne, ev.data(), vmark.data(), (int64_t)sm.size(), sv.data(), sm.data());
      CHECK(s2.n_dof == s.n_dof && s2.col == s.col);
      // the other order on the same mesh, and the two-space pattern both ways
      Space o;
      space_build(o, kind == P1_TRI ? P2_TRI : P1_TRI, nv, coords.data(), ne, ev.data(), vmark.data());
      PairPattern pp, qq;
      pair_build(pp, s, o);
      pair_build(qq, o, s);
      CHECK(pp.m == s.n_dof && pp.n == o.n_dof && pp.rp[pp.m] == qq.rp[qq.m]);
    }
    const Csr A = laplacian_on(s);
    for (int lazy = 0; lazy <= 1; lazy++) {
      AmgParams prm;
      prm.lazy = lazy;
      prm.nmpo = 20;
      Hierarchy h;
      amg_build(A, prm, h);
      CHECK(h.A.size() >= 2);
      if (lazy) amg_lazy_refresh(h, A);
      layouts(h);
    }
  }
  std::printf(g_fail ? "host_check: FAILED\n" : "host_check: ok\n");
  return g_fail;
}
