// host.hpp -- host-side structures of the assemble-and-solve path: synthetic
// meshes, FE space (DOF numbering), the CSR pattern with its element gather
// map, the AMG hierarchy (coarsening stays on the host and is uploaded once),
// multicolouring and the sliced-ELL device layout.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace afe {

#ifndef AFE_KIND_ENUM
#define AFE_KIND_ENUM
enum { P1_TRI = 1, P2_TRI = 2, P1_TET = 3 };
#endif
inline int kind_tdim(int k) { return k == P1_TET ? 3 : 2; }
inline int kind_nloc(int k) { return k == P1_TRI ? 3 : (k == P2_TRI ? 6 : 4); }
inline int kind_nvtx(int k) { return k == P1_TET ? 4 : 3; }

struct Error {
  int code;
  std::string msg;
};

// splitmix64: the seeded stream behind every synthetic input
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

void mesh_square(int n, double jitter, uint64_t seed, double *coords,
                 int32_t *ev, int32_t *vmark);
void mesh_cube(int n, double jitter, uint64_t seed, double *coords, int32_t *ev,
               int32_t *vmark);

// CSR with the reference's layout: square matrices keep the diagonal first in
// each row, the remaining columns ascending; rectangular ones fully ascending.
struct Csr {
  int64_t m = 0, n = 0;
  std::vector<int64_t> rp;
  std::vector<int32_t> col;
  std::vector<double> val;
  int64_t nnz() const { return rp.empty() ? 0 : rp[m]; }
};

struct Space {
  int kind = 0, tdim = 0, n_loc = 0, n_vtx = 0;
  int64_t n_vert = 0, n_elem = 0, n_dof = 0;
  std::vector<double> coords;
  std::vector<int32_t> ev, vmark;
  std::vector<int32_t> elem_dof, dof_mark;
  std::vector<double> dof_point;
  // CSR pattern (reference numbering)
  std::vector<int64_t> rp;
  std::vector<int32_t> col;
  int max_row = 0;
  // element -> CSR gather map: for row i, the adjacent elements in ascending
  // order as e*8 + local index, and per entry n_loc slot offsets in row i
  std::vector<int64_t> adj_ptr;
  std::vector<int32_t> adj;
  std::vector<uint8_t> slot;
};

// FEMSpace::buildDof first-touch numbering + pattern + gather map.  With a
// side list (n_side vertex pairs and their boundary marks, as the mesh files
// give them), a P2 edge DOF takes the mark of its side; without one, an edge
// lying in a single element is marked 1.
void space_build(Space &s, int kind, int64_t n_vert, const double *coords,
                 int64_t n_elem, const int32_t *ev, const int32_t *vmark,
                 int64_t n_side = 0, const int32_t *side_vert = nullptr,
                 const int32_t *side_mark = nullptr);

// BilinearOperator::buildSparsityPattern with two FE spaces on one mesh
// (BilinearOperator.templates.h:64-111 and buildDofInfo :115-172, the
// same-mesh branch): row i (a DOF of space 0) couples with every DOF of
// space 1 on the elements holding i; the compressed pattern sorts each row
// ascending, after the diagonal when the pattern is square.
struct PairPattern {
  int64_t m = 0, n = 0;
  int n1 = 0;
  std::vector<int64_t> rp;
  std::vector<int32_t> col;
  std::vector<uint8_t> slot; // slot[p * n1 + b]: position in row i of column b of adjacency entry p of i
  int64_t seg_max = 0;       // largest 256-row value segment (gather LDS)
  int max_coupling = 0;      // nMaxCouplingDof
};
void pair_build(PairPattern &pp, const Space &s0, const Space &s1);

struct AmgParams {
  double alpha = 0.25;
  unsigned nmpo = 50;
  double nmpsd = 0.382;
  int lazy = 0;
};

// host hierarchy: A[0] is a copy of the fine matrix, P[l]: n_{l+1} x n_l,
// Pt[l] its transpose (the reference keeps it for lazy mode)
struct Hierarchy {
  AmgParams prm;
  std::vector<Csr> A, P, Pt;
  int full = 1;
  std::vector<double> minv; // coarsest dense inverse, row-major
};

void amg_build(const Csr &A0, const AmgParams &prm, Hierarchy &h);
void amg_lazy_refresh(Hierarchy &h, const Csr &A0);
// the coarsest level's dense inverse from its current values (h.minv)
void coarse_inverse(Hierarchy &h);
Csr galerkin(const Csr &P, const Csr &A, const Csr &Pt);
Csr transpose(const Csr &a);
int dense_inverse(int n, double *a);

// ---- multicolour sliced-ELL (device layout) -------------------------------
// Rows are permuted so that each colour is a contiguous range, sorted by
// decreasing stored length inside the colour; slices never straddle a colour.
// A slice is one wavefront: with G = 2^gshift lanes per row it holds
// SELL_C / G consecutive rows, and stored entry j of its row l lives at
// off + (j / G)*SELL_C + l*G + j % G, so lane l*G + k walks entries k, k+G,
// k+2G, ... of row l at off + step*SELL_C + lane (width = steps).  G = 1 is
// the plain one-lane-per-row sliced ELL.  Padding uses the row itself as
// column with value 0.
constexpr int SELL_C = 64;
constexpr int MAX_ROW_LANE_SHIFT = 4; // at most 16 lanes per row

struct SliceInfo {
  int32_t row0, nrow, width, off;
};

struct Sell {
  int64_t n_rows = 0, n_cols = 0;
  std::vector<SliceInfo> slices;
  std::vector<int32_t> group_slice; // slices of group g: [group_slice[g], group_slice[g+1])
  std::vector<int32_t> group_row;   // rows of group g
  std::vector<int32_t> cols;
  std::vector<double> vals;
  std::vector<double> diag; // only for square matrices stored without diagonal
  int gshift = 0;           // 2^gshift lanes per row
  int64_t stored() const { return (int64_t)cols.size(); }
};

// lanes per row (as a shift) of a matrix whose rows are launched in groups
// (colours) of rows/groups rows on average with `entries` stored entries in
// all: rows are split over more lanes while every lane keeps at least 4
// entries and a group stays under 2^16 lanes (the small levels of the
// V-cycle, whose colours are a few thousand rows of long rows, are latency
// bound with one lane per row), and further, while every lane keeps an
// entry, as long as a group's rows fill less than one wavefront; max_shift
// caps it
int row_lane_shift(int64_t rows, int64_t groups, int64_t entries, int max_shift);

// colouring of the symmetrised graph of the stored off-diagonal entries
// (explicit zeros included): DSatur for matrices of at most dsatur_max_rows
// rows, greedy first-fit in index order above (linear time).  The fine P1/P2
// matrices get the same colour count either way (C2's 4.2M rows: 6 and 6,
// DSatur taking 4.3 s against 0.7 s), so level 0 uses DSatur only up to
// DSATUR_MAX_ROWS; the Galerkin coarse levels gain much more (the 3D level
// of 537K rows: 9 colours against 20 in 1.1 s against 0.4 s; the 2D P2
// level of 481K rows: 6 against 7) and use it up to DSATUR_MAX_ROWS_COARSE.
constexpr int64_t DSATUR_MAX_ROWS = 1 << 18;
constexpr int64_t DSATUR_MAX_ROWS_COARSE = 1 << 22;
int colour_rows(const Csr &a, std::vector<int32_t> &colour, int64_t dsatur_max_rows = DSATUR_MAX_ROWS);

// perm: new -> old row order; col_map: old column -> new column
// groups: row-count per group in new order (slices never cross groups)
// src_map (optional): position of every CSR entry in the layout, -1 for a
// split diagonal; diag_src (optional): CSR index of each stored row's diagonal
// gshift: 2^gshift lanes per row
void build_sell(const Csr &a, const std::vector<int32_t> &perm,
                const std::vector<int32_t> &col_map,
                const std::vector<int32_t> &group_rows, bool split_diag,
                Sell &out, std::vector<int32_t> *src_map = nullptr,
                std::vector<int32_t> *diag_src = nullptr, int gshift = 0);

// host threads of the parallel loops started by the calling thread
void set_host_threads(int n);

} // namespace afe
