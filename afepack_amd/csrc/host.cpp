// host.cpp -- host-side algorithms of the assemble-and-solve path.
//
// The DOF numbering, CSR pattern and AMG coarsening are restatements of the
// reference algorithms (citations below, relative to the AFEPack tree); they
// produce the same patterns and, row by row, the same floating-point
// accumulation order, but use flat arrays, binary searches and OpenMP over
// independent rows instead of std::list scans.
#include "host.hpp"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <queue>
#include <tuple>
#include <stdexcept>

namespace afe {

// ===========================================================================
// synthetic meshes
// ===========================================================================
void mesh_square(int n, double jitter, uint64_t seed, double *coords,
                 int32_t *ev, int32_t *vmark) {
  const int64_t nv1 = n + 1;
  const double h = 1.0 / n;
  for (int64_t j = 0; j <= n; j++)
    for (int64_t i = 0; i <= n; i++) {
      int64_t v = j * nv1 + i;
      coords[2 * v] = (double)i / n;
      coords[2 * v + 1] = (double)j / n;
      vmark[v] = (i == 0 || i == n || j == 0 || j == n) ? 1 : 0;
    }
  if (jitter > 0) {
    Rng rng(seed);
    for (int64_t v = 0; v < nv1 * nv1; v++) {
      if (vmark[v]) continue;
      double dx = (2.0 * rng.uniform() - 1.0) * jitter * h;
      double dy = (2.0 * rng.uniform() - 1.0) * jitter * h;
      coords[2 * v] += dx;
      coords[2 * v + 1] += dy;
    }
  }
  // squares row by row, each split along its (n00, n11) diagonal, as the
  // structured writer of poisson_lshape_stretched_afepack.cpp:262-286
  int64_t e = 0;
  for (int64_t j = 0; j < n; j++)
    for (int64_t i = 0; i < n; i++) {
      int32_t n00 = (int32_t)(j * nv1 + i), n10 = n00 + 1;
      int32_t n01 = (int32_t)(n00 + nv1), n11 = n01 + 1;
      int32_t *t = ev + 6 * e;
      t[0] = n00; t[1] = n10; t[2] = n11;
      t[3] = n00; t[4] = n11; t[5] = n01;
      e++;
    }
}

void mesh_cube(int n, double jitter, uint64_t seed, double *coords, int32_t *ev,
               int32_t *vmark) {
  const int64_t nv1 = n + 1;
  const double h = 1.0 / n;
  // vertex numbering and the 6-tetrahedron split of tools/tricube.cpp:54-90
  // (x slowest, z fastest)
  for (int64_t i = 0; i <= n; i++)
    for (int64_t j = 0; j <= n; j++)
      for (int64_t k = 0; k <= n; k++) {
        int64_t v = (i * nv1 + j) * nv1 + k;
        coords[3 * v] = (double)i / n;
        coords[3 * v + 1] = (double)j / n;
        coords[3 * v + 2] = (double)k / n;
        vmark[v] = (i == 0 || i == n || j == 0 || j == n || k == 0 || k == n) ? 1 : 0;
      }
  if (jitter > 0) {
    Rng rng(seed);
    for (int64_t v = 0; v < nv1 * nv1 * nv1; v++) {
      if (vmark[v]) continue;
      for (int d = 0; d < 3; d++) coords[3 * v + d] += (2.0 * rng.uniform() - 1.0) * jitter * h;
    }
  }
  static const int8_t split[6][4][3] = {
      {{0, 0, 0}, {0, 0, 1}, {1, 0, 0}, {0, 1, 0}},
      {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}},
      {{0, 1, 1}, {0, 1, 0}, {1, 0, 1}, {0, 0, 1}},
      {{1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {1, 0, 1}},
      {{0, 1, 1}, {1, 1, 0}, {1, 0, 1}, {0, 1, 0}},
      {{1, 1, 1}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}}};
  int64_t e = 0;
  for (int64_t i = 0; i < n; i++)
    for (int64_t j = 0; j < n; j++)
      for (int64_t k = 0; k < n; k++)
        for (int t = 0; t < 6; t++, e++)
          for (int q = 0; q < 4; q++)
            ev[4 * e + q] = (int32_t)(((i + split[t][q][0]) * nv1 + (j + split[t][q][1])) * nv1 +
                                      (k + split[t][q][2]));
}

// ===========================================================================
// FE space
// ===========================================================================
// FEMSpace::buildDof (FEMSpace.templates.h:705-880, one thread): elements in
// order, dimensions 0..TDIM, template sub-geometries in template order, DOFs
// on first touch; triangle edge k is opposite vertex k (triangle.tmp_geo:21-30).
// The pattern is BilinearOperator::buildSparsityPattern
// (BilinearOperator.templates.h:64-111): the diagonal and every DOF sharing an
// element, diagonal first, the rest ascending.
void space_build(Space &s, int kind, int64_t n_vert, const double *coords,
                 int64_t n_elem, const int32_t *ev, const int32_t *vmark,
                 int64_t n_side, const int32_t *side_vert, const int32_t *side_mark) {
  s.kind = kind;
  s.tdim = kind_tdim(kind);
  s.n_loc = kind_nloc(kind);
  s.n_vtx = kind_nvtx(kind);
  s.n_vert = n_vert;
  s.n_elem = n_elem;
  const int dim = s.tdim, nv = s.n_vtx, nl = s.n_loc;
  s.coords.assign(coords, coords + n_vert * dim);
  s.ev.assign(ev, ev + n_elem * nv);
  s.vmark.assign(vmark, vmark + n_vert);
  for (int64_t i = 0; i < n_elem * nv; i++)
    if (ev[i] < 0 || ev[i] >= n_vert) throw Error{-1, "element vertex index out of range"};

  // mesh edges of P2 triangles: per smaller-vertex lists
  std::vector<int64_t> edge_of, edge_cnt;
  if (kind == P2_TRI) {
    std::vector<int64_t> off(n_vert + 1, 0);
    for (int64_t e = 0; e < n_elem; e++)
      for (int k = 0; k < 3; k++) {
        int32_t p = ev[3 * e + (k + 1) % 3], q = ev[3 * e + (k + 2) % 3];
        off[std::min(p, q) + 1]++;
      }
    for (int64_t v = 0; v < n_vert; v++) off[v + 1] += off[v];
    std::vector<int64_t> rec(off[n_vert]);
    std::vector<int64_t> fill(off.begin(), off.end() - 1);
    for (int64_t e = 0; e < n_elem; e++)
      for (int k = 0; k < 3; k++) {
        int32_t p = ev[3 * e + (k + 1) % 3], q = ev[3 * e + (k + 2) % 3];
        int32_t a = std::min(p, q), b = std::max(p, q);
        rec[fill[a]++] = ((int64_t)b << 32) | (e * 3 + k);
      }
    edge_of.assign(n_elem * 3, -1);
    int64_t n_edge = 0;
    for (int64_t v = 0; v < n_vert; v++) {
      std::sort(rec.begin() + off[v], rec.begin() + off[v + 1]);
      for (int64_t r = off[v]; r < off[v + 1]; r++) {
        if (r > off[v] && (rec[r] >> 32) != (rec[r - 1] >> 32)) n_edge++;
        edge_of[rec[r] & 0xffffffffll] = n_edge;
      }
      if (off[v + 1] > off[v]) n_edge++;
    }
    edge_cnt.assign(n_edge, 0);
    for (int64_t i = 0; i < n_elem * 3; i++) edge_cnt[edge_of[i]]++;
  }
  // buildDofBoundaryMark (FEMSpace.templates.h:689-701) gives an edge DOF the
  // boundary mark of its mesh side: look each element edge up in the side
  // list by its vertex pair
  std::vector<int32_t> edge_mark;
  if (kind == P2_TRI && n_side > 0) {
    if (!side_vert || !side_mark) throw Error{-1, "side list without vertices or marks"};
    std::vector<std::pair<int64_t, int32_t>> key(n_side);
    for (int64_t i = 0; i < n_side; i++) {
      const int32_t a = side_vert[2 * i], b = side_vert[2 * i + 1];
      if (a < 0 || b < 0 || a >= n_vert || b >= n_vert) throw Error{-1, "side vertex index out of range"};
      key[i] = {((int64_t)std::min(a, b) << 32) | std::max(a, b), side_mark[i]};
    }
    std::sort(key.begin(), key.end(),
              [](const std::pair<int64_t, int32_t> &x, const std::pair<int64_t, int32_t> &y) { return x.first < y.first; });
    edge_mark.assign(edge_cnt.size(), 0);
    for (int64_t e = 0; e < n_elem; e++)
      for (int k = 0; k < 3; k++) {
        const int32_t p = ev[3 * e + (k + 1) % 3], q = ev[3 * e + (k + 2) % 3];
        const int64_t kk = ((int64_t)std::min(p, q) << 32) | std::max(p, q);
        auto it = std::lower_bound(key.begin(), key.end(), kk,
                                   [](const std::pair<int64_t, int32_t> &x, int64_t v) { return x.first < v; });
        if (it == key.end() || it->first != kk) throw Error{-1, "an element edge is missing from the side list"};
        edge_mark[edge_of[3 * e + k]] = it->second;
      }
  }

  // first-touch numbering
  s.elem_dof.assign(n_elem * nl, -1);
  std::vector<int32_t> vdof(n_vert, -1), edof(edge_cnt.size(), -1);
  int32_t n_dof = 0;
  for (int64_t e = 0; e < n_elem; e++) {
    for (int j = 0; j < nv; j++) {
      int32_t v = ev[nv * e + j];
      if (vdof[v] < 0) vdof[v] = n_dof++;
      s.elem_dof[nl * e + j] = vdof[v];
    }
    if (kind == P2_TRI)
      for (int j = 0; j < 3; j++) {
        int64_t g = edge_of[3 * e + j];
        if (edof[g] < 0) edof[g] = n_dof++;
        s.elem_dof[nl * e + 3 + j] = edof[g];
      }
  }
  s.n_dof = n_dof;
  // boundary marks (buildDofBoundaryMark, FEMSpace.templates.h:689-701) and
  // interpolation points (vertex, or edge midpoint as local_to_global of the
  // template interpolation point)
  s.dof_mark.assign(n_dof, 0);
  s.dof_point.assign((size_t)n_dof * dim, 0.0);
  for (int64_t e = 0; e < n_elem; e++) {
    for (int j = 0; j < nv; j++) {
      int32_t v = ev[nv * e + j], d = s.elem_dof[nl * e + j];
      s.dof_mark[d] = vmark[v];
      for (int c = 0; c < dim; c++) s.dof_point[(size_t)d * dim + c] = coords[(size_t)v * dim + c];
    }
    if (kind == P2_TRI)
      for (int j = 0; j < 3; j++) {
        int32_t d = s.elem_dof[nl * e + 3 + j];
        s.dof_mark[d] = edge_mark.empty() ? (edge_cnt[edge_of[3 * e + j]] == 1 ? 1 : 0) : edge_mark[edge_of[3 * e + j]];
        int32_t p = ev[3 * e + (j + 1) % 3], q = ev[3 * e + (j + 2) % 3];
        for (int c = 0; c < 2; c++)
          s.dof_point[(size_t)d * 2 + c] = 0.5 * coords[(size_t)p * 2 + c] + 0.5 * coords[(size_t)q * 2 + c];
      }
  }

  // DOF -> element adjacency, ascending element order
  s.adj_ptr.assign(n_dof + 1, 0);
  for (int64_t i = 0; i < n_elem * nl; i++) s.adj_ptr[s.elem_dof[i] + 1]++;
  for (int64_t d = 0; d < n_dof; d++) s.adj_ptr[d + 1] += s.adj_ptr[d];
  s.adj.assign(s.adj_ptr[n_dof], 0);
  {
    std::vector<int64_t> fill(s.adj_ptr.begin(), s.adj_ptr.end() - 1);
    for (int64_t e = 0; e < n_elem; e++)
      for (int a = 0; a < nl; a++) s.adj[fill[s.elem_dof[nl * e + a]]++] = (int32_t)(e * 8 + a);
  }

  // pattern rows: diagonal, then the sorted distinct neighbours
  s.rp.assign(n_dof + 1, 0);
  std::vector<int32_t> rowlen(n_dof);
  const int64_t CH = 4096;
  const int64_t nch = (n_dof + CH - 1) / CH;
  std::vector<std::vector<int32_t>> chunk_cols(nch);
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t c = 0; c < nch; c++) {
    std::vector<int32_t> buf;
    // seen[d] == i: d already listed for row i (duplicates are dropped
    // before the sort, which then orders ~15 distinct columns, not ~100)
    std::vector<int32_t> seen(n_dof, -1);
    std::vector<int32_t> &out = chunk_cols[c];
    for (int64_t i = c * CH; i < std::min<int64_t>(n_dof, (c + 1) * CH); i++) {
      buf.clear();
      seen[i] = (int32_t)i;
      for (int64_t p = s.adj_ptr[i]; p < s.adj_ptr[i + 1]; p++) {
        int64_t e = s.adj[p] >> 3;
        for (int b = 0; b < nl; b++) {
          int32_t d = s.elem_dof[nl * e + b];
          if (seen[d] != i) {
            seen[d] = (int32_t)i;
            buf.push_back(d);
          }
        }
      }
      std::sort(buf.begin(), buf.end());
      rowlen[i] = (int32_t)buf.size() + 1;
      out.push_back((int32_t)i);
      out.insert(out.end(), buf.begin(), buf.end());
    }
  }
  for (int64_t i = 0; i < n_dof; i++) s.rp[i + 1] = s.rp[i] + rowlen[i];
  s.col.resize(s.rp[n_dof]);
  {
    int64_t pos = 0;
    for (int64_t c = 0; c < nch; c++) {
      std::copy(chunk_cols[c].begin(), chunk_cols[c].end(), s.col.begin() + pos);
      pos += (int64_t)chunk_cols[c].size();
      std::vector<int32_t>().swap(chunk_cols[c]);
    }
  }
  s.max_row = 0;
  for (int64_t i = 0; i < n_dof; i++) s.max_row = std::max(s.max_row, rowlen[i]);
  if (s.max_row > 255) throw Error{-1, "row longer than 255 entries: unsupported gather map"};

  // gather map: slot offset of (row i, column elem_dof[e][b]) inside row i,
  // looked up in a per-thread column -> offset table filled from row i
  s.slot.assign((size_t)s.adj.size() * nl, 0);
#pragma omp parallel
  {
    std::vector<uint8_t> at(n_dof, 0);
#pragma omp for schedule(static, 4096)
    for (int64_t i = 0; i < n_dof; i++) {
      const int32_t *c0 = s.col.data() + s.rp[i];
      const int64_t len = s.rp[i + 1] - s.rp[i];
      for (int64_t o = 0; o < len; o++) at[c0[o]] = (uint8_t)o;
      for (int64_t p = s.adj_ptr[i]; p < s.adj_ptr[i + 1]; p++) {
        const int64_t e = s.adj[p] >> 3;
        for (int b = 0; b < nl; b++) s.slot[(size_t)p * nl + b] = at[s.elem_dof[nl * e + b]];
      }
    }
  }
}

void pair_build(PairPattern &pp, const Space &s0, const Space &s1) {
  // BilinearOperator::buildSparsityPattern, two spaces on one mesh
  // (BilinearOperator.templates.h:64-95): element e couples every space-0
  // DOF of e with every space-1 DOF of e; SparsityPattern::reinit/compress
  // (SparsityPattern.h:16-70,193-235) keeps a square pattern's diagonal
  // first in each row (present even where no element couples it) and sorts
  // the other columns ascending
  if (s0.n_elem != s1.n_elem || s0.ev != s1.ev || s0.coords != s1.coords)
    throw Error{-1, "the two spaces must be built on the same mesh"};
  const int n0 = s0.n_loc, n1 = s1.n_loc;
  pp.m = s0.n_dof;
  pp.n = s1.n_dof;
  pp.n1 = n1;
  const bool diag_first = pp.m == pp.n;
  pp.rp.assign(pp.m + 1, 0);
  pp.col.clear();
  pp.slot.assign(s0.adj.size() * n1, 0);
  std::vector<int32_t> row;
  for (int64_t i = 0; i < pp.m; i++) {
    row.clear();
    for (int64_t p = s0.adj_ptr[i]; p < s0.adj_ptr[i + 1]; p++) {
      const int64_t e = s0.adj[p] >> 3;
      for (int b = 0; b < n1; b++) row.push_back(s1.elem_dof[e * n1 + b]);
    }
    std::sort(row.begin(), row.end());
    row.erase(std::unique(row.begin(), row.end()), row.end());
    if (diag_first) {
      auto it = std::lower_bound(row.begin(), row.end(), (int32_t)i);
      if (it != row.end() && *it == i) row.erase(it);
      row.insert(row.begin(), (int32_t)i);
    }
    if (row.size() > 255) throw Error{-1, "a row of the two-space pattern exceeds 255 entries"};
    const size_t off = diag_first ? 1 : 0; // the sorted part of the row
    for (int64_t p = s0.adj_ptr[i]; p < s0.adj_ptr[i + 1]; p++) {
      const int64_t e = s0.adj[p] >> 3;
      for (int b = 0; b < n1; b++) {
        const int32_t c = s1.elem_dof[e * n1 + b];
        pp.slot[p * n1 + b] = (uint8_t)(diag_first && c == i
                                            ? 0
                                            : std::lower_bound(row.begin() + off, row.end(), c) - row.begin());
      }
    }
    pp.col.insert(pp.col.end(), row.begin(), row.end());
    pp.rp[i + 1] = (int64_t)pp.col.size();
  }
  pp.seg_max = 0;
  for (int64_t r0 = 0; r0 < pp.m; r0 += 256)
    pp.seg_max = std::max(pp.seg_max, pp.rp[std::min(r0 + 256, pp.m)] - pp.rp[r0]);
  // buildDofInfo, same mesh (BilinearOperator.templates.h:131-171): each
  // element adds its space-1 DOF count to the coupling count of each of its
  // space-0 DOFs; the maximum, capped at the column count
  std::vector<int64_t> cnt(pp.m, 0);
  for (int64_t e = 0; e < s0.n_elem; e++)
    for (int a = 0; a < n0; a++) cnt[s0.elem_dof[e * n0 + a]] += n1;
  int64_t mx = 0;
  for (int64_t c : cnt) mx = std::max(mx, c);
  pp.max_coupling = (int)std::min<int64_t>(mx, pp.n);
}

// ===========================================================================
// AMG coarsening (host, uploaded once)
// ===========================================================================
static inline double csr_el(const Csr &a, int64_t i, int64_t j) {
  // SparseMatrix::el (SparseMatrix.h:58-70); rows hold each column once
  const int64_t b = a.rp[i], e = a.rp[i + 1];
  if (b == e) return 0.0;
  if (a.col[b] == j) return a.val[b];
  const int32_t *p = std::lower_bound(a.col.data() + b + 1, a.col.data() + e, (int32_t)j);
  if (p != a.col.data() + e && *p == j) return a.val[p - a.col.data()];
  // unsorted rows (not produced here) fall back to a scan
  for (int64_t k = b; k < e; k++)
    if (a.col[k] == j) return a.val[k];
  return 0.0;
}

Csr transpose(const Csr &a) {
  Csr t;
  t.m = a.n;
  t.n = a.m;
  t.rp.assign(t.m + 1, 0);
  for (int64_t k = 0; k < a.nnz(); k++) t.rp[a.col[k] + 1]++;
  for (int64_t i = 0; i < t.m; i++) t.rp[i + 1] += t.rp[i];
  t.col.resize(a.nnz());
  t.val.resize(a.nnz());
  std::vector<int64_t> fill(t.rp.begin(), t.rp.end() - 1);
  for (int64_t i = 0; i < a.m; i++)
    for (int64_t k = a.rp[i]; k < a.rp[i + 1]; k++) {
      int64_t p = fill[a.col[k]]++;
      t.col[p] = (int32_t)i;
      t.val[p] = a.val[k];
    }
  return t;
}

// AMGSolver::getPMPt (src/AMGSolver.cpp:288-358): per coarse row, columns in
// discovery order over P row x A row x P^T row with the diagonal first, then
// compressed (diagonal first, rest ascending); each value accumulates
// (P_ia * A_ab) * Pt_bc in the same loop order as the reference.
Csr galerkin(const Csr &P, const Csr &A, const Csr &Pt) {
  const int64_t nc = P.m;
  Csr R;
  R.m = R.n = nc;
  R.rp.assign(nc + 1, 0);
  const int64_t CH = 2048;
  const int64_t nch = (nc + CH - 1) / CH;
  std::vector<std::vector<int32_t>> ccol(nch);
  std::vector<std::vector<double>> cval(nch);
  std::vector<int64_t> rowlen(nc);
#pragma omp parallel
  {
    std::vector<char> flag(nc, 1);
    std::vector<double> row_entry(nc, 0.0);
    std::vector<int32_t> idx;
#pragma omp for schedule(dynamic, 1)
    for (int64_t c = 0; c < nch; c++) {
      for (int64_t i = c * CH; i < std::min(nc, (c + 1) * CH); i++) {
        idx.clear();
        idx.push_back((int32_t)i);
        flag[i] = 0;
        for (int64_t j = P.rp[i]; j < P.rp[i + 1]; j++) {
          const int64_t a = P.col[j];
          const double pa = P.val[j];
          for (int64_t k = A.rp[a]; k < A.rp[a + 1]; k++) {
            const int64_t b = A.col[k];
            const double pab = pa * A.val[k];
            for (int64_t l = Pt.rp[b]; l < Pt.rp[b + 1]; l++) {
              const int32_t cc = Pt.col[l];
              if (flag[cc]) {
                idx.push_back(cc);
                flag[cc] = 0;
              }
              row_entry[cc] += pab * Pt.val[l];
            }
          }
        }
        std::sort(idx.begin() + 1, idx.end());
        rowlen[i] = (int64_t)idx.size();
        for (int32_t cc : idx) {
          ccol[c].push_back(cc);
          cval[c].push_back(row_entry[cc]);
          row_entry[cc] = 0.0;
          flag[cc] = 1;
        }
      }
    }
  }
  for (int64_t i = 0; i < nc; i++) R.rp[i + 1] = R.rp[i] + rowlen[i];
  R.col.resize(R.rp[nc]);
  R.val.resize(R.rp[nc]);
  int64_t pos = 0;
  for (int64_t c = 0; c < nch; c++) {
    std::copy(ccol[c].begin(), ccol[c].end(), R.col.begin() + pos);
    std::copy(cval[c].begin(), cval[c].end(), R.val.begin() + pos);
    pos += (int64_t)ccol[c].size();
  }
  return R;
}

// AMGSolver::Project (src/AMGSolver.cpp:18-226).  Steps 1-2 (C/F split and
// the F-F fix) are sequential as in the reference; the interpolation rows are
// independent and run in parallel.  P^T row i is assembled first (row i of
// the fine matrix writes only there) and P is its transpose, which matches
// the reference's twin accumulation into P and Pt bit for bit.
static void project(const Csr &M, double alpha, Csr &P, Csr &Pt) {
  const int64_t n = M.m;
  const int64_t *rp = M.rp.data();
  const int32_t *col = M.col.data();
  const double *v = M.val.data();
  std::vector<int8_t> flag(n, 0);
  std::vector<double> beta(n, 0.0);
  int64_t n_core = 0;
  for (int64_t i = 0; i < n; i++) {
    double b = 0.0;
    for (int64_t j = rp[i] + 1; j < rp[i + 1]; j++) b = std::max(std::fabs(v[j]), b);
    beta[i] = b * alpha;
    if (flag[i] != 0) continue;
    for (int64_t j = rp[i] + 1; j < rp[i + 1]; j++) {
      const int64_t k = col[j];
      if (flag[k] == -1) continue;
      if (std::fabs(v[j]) < beta[i]) continue;
      flag[k] = 1;
    }
    flag[i] = -1;
    n_core++;
  }
  std::vector<int64_t> mark(n, -1);
  for (int64_t i = 0; i < n; i++) {
    if (flag[i] != 1) continue;
    for (int64_t j = rp[i] + 1; j < rp[i + 1]; j++)
      if (!(std::fabs(v[j]) < beta[i]) && flag[col[j]] == -1) mark[col[j]] = i;
    bool is_FF = false;
    for (int64_t j = rp[i] + 1; j < rp[i + 1]; j++) {
      const int64_t k = col[j];
      if (flag[k] == -1 || k < i || std::fabs(v[j]) < beta[i]) continue;
      is_FF = true;
      for (int64_t j1 = rp[k] + 1; j1 < rp[k + 1]; j1++) {
        const int64_t k1 = col[j1];
        if (flag[k1] != -1) continue;
        if (mark[k1] == i) {
          is_FF = false;
          break;
        }
      }
      if (is_FF) break;
    }
    if (!is_FF) continue;
    flag[i] = -1;
    n_core++;
  }
  std::vector<int32_t> core_point(n_core);
  std::vector<int32_t> core_at(n, -1);
  for (int64_t i = 0, c = 0; i < n; i++)
    if (flag[i] == -1) {
      core_point[c] = (int32_t)i;
      core_at[i] = (int32_t)c++;
    }
  // P^T pattern: row l lists (ascending) the coarse points c whose core row
  // has l as a non-C neighbour, and its own coarse index when l is a C point
  Pt = Csr();
  Pt.m = n;
  Pt.n = n_core;
  Pt.rp.assign(n + 1, 0);
  for (int64_t c = 0; c < n_core; c++) {
    const int64_t k = core_point[c];
    Pt.rp[k + 1]++;
    for (int64_t j = rp[k] + 1; j < rp[k + 1]; j++)
      if (flag[col[j]] != -1) Pt.rp[col[j] + 1]++;
  }
  for (int64_t i = 0; i < n; i++) Pt.rp[i + 1] += Pt.rp[i];
  Pt.col.resize(Pt.rp[n]);
  Pt.val.assign(Pt.rp[n], 0.0);
  {
    std::vector<int64_t> fill(Pt.rp.begin(), Pt.rp.end() - 1);
    for (int64_t c = 0; c < n_core; c++) {
      const int64_t k = core_point[c];
      Pt.col[fill[k]++] = (int32_t)c;
      for (int64_t j = rp[k] + 1; j < rp[k + 1]; j++)
        if (flag[col[j]] != -1) Pt.col[fill[col[j]]++] = (int32_t)c;
    }
  }
  const int64_t *Trp = Pt.rp.data();
  const int32_t *Tc = Pt.col.data();
  auto pt_add = [&](int64_t i, int32_t c, double w) {
    const int32_t *p = std::lower_bound(Tc + Trp[i], Tc + Trp[i + 1], c);
    if (p != Tc + Trp[i + 1] && *p == c) Pt.val[p - Tc] += w;
  };
  auto in_row = [&](int64_t i, int32_t c) {
    return std::binary_search(Tc + Trp[i], Tc + Trp[i + 1], c);
  };
#pragma omp parallel for schedule(dynamic, 2048)
  for (int64_t i = 0; i < n; i++) {
    if (core_at[i] >= 0) {
      pt_add(i, core_at[i], 1.0);
      continue;
    }
    double aii = v[rp[i]];
    for (int64_t j = rp[i] + 1; j < rp[i + 1]; j++)
      if (std::fabs(v[j]) < beta[i]) aii += v[j];
    for (int64_t j = rp[i] + 1; j < rp[i + 1]; j++) {
      const int64_t l = col[j];
      if (std::fabs(v[j]) < beta[i]) continue;
      if (core_at[l] >= 0) {
        pt_add(i, core_at[l], std::fabs(v[j] / aii));
      } else {
        double d = 0.0;
        for (int64_t i1 = Trp[l]; i1 < Trp[l + 1]; i1++) {
          const int32_t c1 = Tc[i1];
          if (in_row(i, c1)) d += csr_el(M, l, core_point[c1]);
        }
        if (std::fabs(d * aii) < 1e-16) continue;
        const double ail = csr_el(M, i, l);
        for (int64_t i1 = Trp[l]; i1 < Trp[l + 1]; i1++) {
          const int32_t c1 = Tc[i1];
          if (in_row(i, c1))
            pt_add(i, c1, std::fabs(ail * csr_el(M, l, core_point[c1]) / (d * aii)));
        }
      }
    }
  }
  P = transpose(Pt);
}

// SparsityPattern::max_entries_per_row (SparsityPattern.h:239-249): the last
// row is not part of the maximum
static int64_t max_entries_per_row(const Csr &a) {
  int64_t m = 0;
  for (int64_t i = 1; i < a.m; i++) m = std::max(m, a.rp[i] - a.rp[i - 1]);
  return m;
}

// AMGSolver::lazyProject (src/AMGSolver.cpp:235-284), including the row
// capacity max_entries_per_row(M) of both P and P^T (adds beyond it are
// dropped, SparsityPattern.h:251-262)
static void lazy_project(const Csr &M, Csr &P, Csr &Pt) {
  const int64_t n = M.m;
  std::vector<uint32_t> counter(n, 0);
  std::vector<int32_t> core_point;
  for (int64_t i = 0; i < n; i++) {
    if (counter[i] > 0) continue;
    for (int64_t j = M.rp[i] + 1; j < M.rp[i + 1]; j++) counter[M.col[j]]++;
    core_point.push_back((int32_t)i);
  }
  const int64_t nc = (int64_t)core_point.size();
  const int64_t cap = max_entries_per_row(M);
  // P rows: core k then its neighbours (add order; a repeated column adds to
  // its entry), capped, then sorted by column
  P = Csr();
  P.m = nc;
  P.n = M.n;
  P.rp.assign(nc + 1, 0);
  P.col.reserve(M.nnz());
  P.val.reserve(M.nnz());
  std::vector<std::pair<int32_t, double>> r;
  for (int64_t c = 0; c < nc; c++) {
    const int64_t k = core_point[c];
    r.clear();
    auto add = [&](int32_t j, double w) {
      for (auto &e : r)
        if (e.first == j) {
          e.second += w;
          return;
        }
      if ((int64_t)r.size() < cap) r.push_back({j, w});
    };
    add((int32_t)k, 1.0);
    for (int64_t j = M.rp[k] + 1; j < M.rp[k + 1]; j++) add(M.col[j], 1.0 / counter[M.col[j]]);
    std::sort(r.begin(), r.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
    for (auto &e : r) {
      P.col.push_back(e.first);
      P.val.push_back(e.second);
    }
    P.rp[c + 1] = (int64_t)P.col.size();
  }
  // P^T rows receive entries in core order, capped likewise: a row keeps the
  // first cap cores that reach it; one core reaches a row once (a repeated
  // column of the core's row adds to that entry, which is the row's last)
  Pt = Csr();
  Pt.m = M.n;
  Pt.n = nc;
  std::vector<int64_t> cnt(M.n, 0);
  std::vector<int32_t> last(M.n, -1);
  auto reach = [&](int64_t l, int32_t c) {
    if (last[l] != c) {
      last[l] = c;
      if (cnt[l] < cap) cnt[l]++;
    }
  };
  for (int64_t c = 0; c < nc; c++) {
    const int64_t k = core_point[c];
    reach(k, (int32_t)c);
    for (int64_t j = M.rp[k] + 1; j < M.rp[k + 1]; j++) reach(M.col[j], (int32_t)c);
  }
  Pt.rp.assign(M.n + 1, 0);
  for (int64_t l = 0; l < M.n; l++) Pt.rp[l + 1] = Pt.rp[l] + cnt[l];
  Pt.col.assign(Pt.rp[M.n], 0);
  Pt.val.assign(Pt.rp[M.n], 0.0);
  std::vector<int64_t> fill(Pt.rp.begin(), Pt.rp.end() - 1);
  auto put = [&](int64_t l, int32_t c, double w) {
    if (fill[l] > Pt.rp[l] && Pt.col[fill[l] - 1] == c) {
      Pt.val[fill[l] - 1] += w;
      return;
    }
    if (fill[l] < Pt.rp[l + 1]) {
      Pt.col[fill[l]] = c;
      Pt.val[fill[l]++] = w;
    }
  };
  for (int64_t c = 0; c < nc; c++) {
    const int64_t k = core_point[c];
    put(k, (int32_t)c, 1.0);
    for (int64_t j = M.rp[k] + 1; j < M.rp[k + 1]; j++) put(M.col[j], (int32_t)c, 1.0 / counter[M.col[j]]);
  }
}

// Gauss-Jordan inverse with partial pivoting (stands in for LAPACKE
// dgetrf + dgetri of DenseMatrix::gauss_jordan, DenseMatrix.h:131-172)
int dense_inverse(int n, double *a) {
  std::vector<double> b((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++) b[(size_t)i * n + i] = 1.0;
  for (int c = 0; c < n; c++) {
    int p = c;
    for (int r = c + 1; r < n; r++)
      if (std::fabs(a[r * n + c]) > std::fabs(a[p * n + c])) p = r;
    if (a[p * n + c] == 0.0) return -1;
    if (p != c)
      for (int j = 0; j < n; j++) {
        std::swap(a[c * n + j], a[p * n + j]);
        std::swap(b[c * n + j], b[p * n + j]);
      }
    const double piv = a[c * n + c];
    for (int j = 0; j < n; j++) {
      a[c * n + j] /= piv;
      b[c * n + j] /= piv;
    }
    // row updates are independent: parallel for large coarse levels
#pragma omp parallel for schedule(static) if (n > 256)
    for (int r = 0; r < n; r++) {
      if (r == c) continue;
      const double f = a[(size_t)r * n + c];
      if (f == 0.0) continue;
      for (int j = 0; j < n; j++) {
        a[(size_t)r * n + j] -= f * a[(size_t)c * n + j];
        b[(size_t)r * n + j] -= f * b[(size_t)c * n + j];
      }
    }
  }
  std::memcpy(a, b.data(), sizeof(double) * n * n);
  return 0;
}

void coarse_inverse(Hierarchy &h) {
  const Csr &Ac = h.A.back();
  const int n = (int)Ac.m;
  h.minv.assign((size_t)n * n, 0.0);
  for (int k = 0; k < n; k++)
    for (int64_t m = Ac.rp[k]; m < Ac.rp[k + 1]; m++) h.minv[(size_t)k * n + Ac.col[m]] = Ac.val[m];
  if (dense_inverse(n, h.minv.data()) != 0) throw Error{-4, "coarsest matrix is singular"};
}

// AMGSolver::reinit / lazyInit (src/AMGSolver.cpp:384-496): project until the
// order is <= nmpo and the float density nnz/m^2 is >= nmpsd; a projection
// that does not reduce the order ends the loop and the coarsest level is then
// solved by diagonal scaling.
void amg_build(const Csr &A0, const AmgParams &prm, Hierarchy &h) {
  h = Hierarchy();
  h.prm = prm;
  h.A.push_back(A0);
  uint32_t k, m;
  k = m = (uint32_t)A0.m;
  float sparse_degree = (float)A0.nnz() / ((float)m * (float)m);
  h.full = 1;
  while (m > prm.nmpo || sparse_degree < prm.nmpsd) {
    Csr P, Pt;
    if (prm.lazy)
      lazy_project(h.A.back(), P, Pt);
    else
      project(h.A.back(), prm.alpha, P, Pt);
    Csr Ac = galerkin(P, h.A.back(), Pt);
    m = (uint32_t)Ac.m;
    if (k == m || h.P.size() >= 60) {
      h.full = 0;
      break;
    }
    k = m;
    sparse_degree = (float)Ac.nnz() / ((float)m * (float)m);
    h.P.push_back(std::move(P));
    h.Pt.push_back(std::move(Pt));
    h.A.push_back(std::move(Ac));
  }
  if (h.full) coarse_inverse(h);
}

// AMGSolver::lazyReinit (src/AMGSolver.cpp:505-545): keep P, recompute the
// Galerkin products and the coarse inverse
void amg_lazy_refresh(Hierarchy &h, const Csr &A0) {
  h.A[0].val = A0.val;
  for (size_t j = 0; j < h.P.size(); j++) h.A[j + 1] = galerkin(h.P[j], h.A[j], h.Pt[j]);
  if (h.full) coarse_inverse(h);
}

// ===========================================================================
// colouring and sliced ELL
// ===========================================================================
int colour_rows(const Csr &a, std::vector<int32_t> &colour, int64_t dsatur_max_rows) {
  const int64_t n = a.m;
  // symmetrised adjacency of the off-diagonal pattern entries (explicit
  // zeros included: the colouring stays valid for any values on the pattern)
  std::vector<int64_t> deg(n + 1, 0);
  for (int64_t i = 0; i < n; i++)
    for (int64_t k = a.rp[i]; k < a.rp[i + 1]; k++)
      if (a.col[k] != i) {
        deg[i + 1]++;
        deg[a.col[k] + 1]++;
      }
  for (int64_t i = 0; i < n; i++) deg[i + 1] += deg[i];
  std::vector<int32_t> adj(deg[n]);
  {
    std::vector<int64_t> fill(deg.begin(), deg.end() - 1);
    for (int64_t i = 0; i < n; i++)
      for (int64_t k = a.rp[i]; k < a.rp[i + 1]; k++)
        if (a.col[k] != i) {
          adj[fill[i]++] = a.col[k];
          adj[fill[a.col[k]]++] = (int32_t)i;
        }
  }
  colour.assign(n, -1);
  if (n <= dsatur_max_rows) {
    // DSatur (Brelaz 1979): colour next the uncoloured row whose neighbours
    // already use the most distinct colours (ties: more neighbours, then the
    // lower index), with the lowest colour its neighbours do not use.  On
    // the coarse levels it needs fewer colours than first-fit in index
    // order, and every colour is one launch per sweep.
    std::vector<uint64_t> used(n, 0); // neighbour colours (at most 64 here)
    std::vector<int32_t> sat(n, 0);
    // heap key: saturation (bits 56..63), degree (32..55), n-1-row (0..31),
    // so the largest key is the most saturated, then the most connected,
    // then the lowest row
    auto key = [&](int64_t i) {
      const uint64_t d = std::min<int64_t>(deg[i + 1] - deg[i], (1 << 24) - 1);
      return (uint64_t)sat[i] << 56 | d << 32 | (uint64_t)(n - 1 - i);
    };
    std::vector<uint64_t> init(n);
    for (int64_t i = 0; i < n; i++) init[i] = key(i);
    std::priority_queue<uint64_t> heap(std::less<uint64_t>(), std::move(init));
    int ncol = 0;
    bool overflow = false;
    while (!heap.empty()) {
      const uint64_t top = heap.top();
      heap.pop();
      const int64_t i = n - 1 - (int64_t)(top & 0xffffffffu);
      if (colour[i] >= 0 || (int32_t)(top >> 56) != sat[i]) continue; // stale entry
      int c = 0;
      while (c < 64 && (used[i] >> c & 1)) c++;
      if (c == 64) {
        overflow = true;
        break;
      }
      colour[i] = c;
      ncol = std::max(ncol, c + 1);
      for (int64_t k = deg[i]; k < deg[i + 1]; k++) {
        const int64_t j = adj[k];
        if (colour[j] >= 0 || (used[j] >> c & 1)) continue;
        used[j] |= uint64_t(1) << c;
        sat[j]++;
        heap.push(key(j));
      }
    }
    if (!overflow) return ncol;
    colour.assign(n, -1); // more than 64 colours: first-fit below
  }
  std::vector<int64_t> stamp(64, -1);
  int ncol = 0;
  for (int64_t i = 0; i < n; i++) {
    for (int64_t k = deg[i]; k < deg[i + 1]; k++) {
      int c = colour[adj[k]];
      if (c >= 0) {
        if (c >= (int)stamp.size()) stamp.resize(2 * c + 2, -1);
        stamp[c] = i;
      }
    }
    int c = 0;
    while (c < (int)stamp.size() && stamp[c] == i) c++;
    if (c >= (int)stamp.size()) stamp.resize(2 * c + 2, -1);
    colour[i] = c;
    ncol = std::max(ncol, c + 1);
  }
  return ncol;
}

int row_lane_shift(int64_t rows, int64_t groups, int64_t entries, int max_shift) {
  if (rows <= 0 || groups <= 0) return 0;
  const int64_t per_group = (rows + groups - 1) / groups;
  const int smax = std::min(max_shift, MAX_ROW_LANE_SHIFT);
  int s = 0;
  while (s < smax && (int64_t(2) << s) * 4 * rows <= entries && (per_group << s) < (int64_t(1) << 16)) s++;
  // a group of a few long rows (the small coarse levels: 2-10 rows per
  // colour) fills a wavefront's 64 lanes with more lanes per row (while
  // every lane keeps at least one entry on average): fewer slices and less
  // padding than 64/G rows per slice would leave
  while (s < smax && (per_group << s) < 64 && (int64_t(1) << s) * 2 * rows <= entries) s++;
  return s;
}

void build_sell(const Csr &a, const std::vector<int32_t> &perm,
                const std::vector<int32_t> &col_map,
                const std::vector<int32_t> &group_rows, bool split_diag,
                Sell &out, std::vector<int32_t> *src_map,
                std::vector<int32_t> *diag_src, int gshift) {
  if (gshift < 0 || gshift > MAX_ROW_LANE_SHIFT) throw Error{-1, "lanes per row out of range"};
  out = Sell();
  out.gshift = gshift;
  const int G = 1 << gshift, R = SELL_C >> gshift; // lanes per row, rows per slice
  out.n_rows = (int64_t)perm.size();
  out.n_cols = a.n;
  if (split_diag) out.diag.assign(out.n_rows, 0.0);
  // every pattern entry is stored (explicit zeros too), so that the layout
  // holds any values on the pattern
  auto stored_len = [&](int64_t o) {
    int64_t c = 0;
    for (int64_t k = a.rp[o]; k < a.rp[o + 1]; k++) {
      if (split_diag && a.col[k] == o) continue;
      c++;
    }
    return c;
  };
  int64_t r = 0, off = 0;
  out.group_slice.push_back(0);
  out.group_row.push_back(0);
  for (int32_t gr : group_rows) {
    const int64_t rend = r + gr;
    for (int64_t r0 = r; r0 < rend; r0 += R) {
      const int32_t nrow = (int32_t)std::min<int64_t>(R, rend - r0);
      int64_t len = 0;
      for (int32_t l = 0; l < nrow; l++) len = std::max(len, stored_len(perm[r0 + l]));
      const int64_t width = (len + G - 1) >> gshift; // steps of SELL_C entries
      out.slices.push_back({(int32_t)r0, nrow, (int32_t)width, (int32_t)off});
      off += width * SELL_C;
    }
    r = rend;
    out.group_slice.push_back((int32_t)out.slices.size());
    out.group_row.push_back((int32_t)r);
  }
  if (off > INT32_MAX) throw Error{-1, "sliced-ELL storage exceeds 2^31 entries"};
  out.cols.assign(off, 0);
  out.vals.assign(off, 0.0);
  if (src_map) src_map->assign(a.nnz(), -1);
  if (diag_src) diag_src->assign(out.n_rows, -1);
#pragma omp parallel for schedule(static, 64)
  for (int64_t s = 0; s < (int64_t)out.slices.size(); s++) {
    const SliceInfo si = out.slices[s];
    for (int32_t l = 0; l < si.nrow; l++) {
      const int64_t nr = si.row0 + l, o = perm[nr];
      const int32_t pad_col = a.m == a.n ? (int32_t)nr : 0;
      int64_t j = 0;
      for (int64_t k = a.rp[o]; k < a.rp[o + 1]; k++) {
        if (split_diag && a.col[k] == o) {
          out.diag[nr] = a.val[k];
          if (diag_src) (*diag_src)[nr] = (int32_t)k;
          continue;
        }
        const int64_t pos = si.off + (j >> gshift) * SELL_C + (l << gshift) + (j & (G - 1));
        out.cols[pos] = col_map[a.col[k]];
        out.vals[pos] = a.val[k];
        if (src_map) (*src_map)[k] = (int32_t)pos;
        j++;
      }
      for (; j < (int64_t)si.width << gshift; j++) {
        const int64_t pos = si.off + (j >> gshift) * SELL_C + (l << gshift) + (j & (G - 1));
        out.cols[pos] = pad_col;
        out.vals[pos] = 0.0;
      }
    }
  }
}

// host threads of the parallel loops started by the calling thread
void set_host_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

} // namespace afe
