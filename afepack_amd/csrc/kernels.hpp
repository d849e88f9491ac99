// kernels.hpp -- gfx950 kernels of the assemble-and-solve path.
//
// Element assembly (FP64, VALU), the element -> CSR gather (LDS accumulation,
// ascending element order, no atomics), right-hand side, Dirichlet
// elimination, error functionals, the multicolour sliced-ELL Gauss-Seidel /
// SpMV family used by the V-cycle (templated on FP64 / FP32), the
// single-workgroup coarse-tail V-cycle and the FP64 CG vector kernels.
// Wavefront = 64 lanes; sliced-ELL slices are 64 rows, one row per lane.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "element_geom.hpp"
#include "host.hpp"

namespace afe {

constexpr int WAVE = 64;
constexpr int WPB = 4;             // wavefronts per block in the SpMV family
constexpr int BLK = WAVE * WPB;    // 256 threads
constexpr int RED_BLOCKS = 1024;   // fixed grid of the reduction kernels
constexpr int TAIL_THREADS = 1024; // single-workgroup coarse tail

// Workgroups are dispatched round-robin over the 8 XCDs; with xcd != 0 the
// logical block index makes consecutive logical blocks share an XCD, so each
// XCD's L2 serves a contiguous (spatially local) range of slices.  gridDim.x
// is a multiple of 8 when xcd != 0.
__device__ inline int logical_block(int xcd) {
  const int b = blockIdx.x;
  return xcd == 1 ? (b & 7) * (int)(gridDim.x >> 3) + (b >> 3) : b;
}

struct FieldDev {
  int n_modes, dim;
  const double *k, *phi, *w;
  double scale;
};

__device__ inline double field_g(const FieldDev &f, const double *x, double *grad) {
  double g = 0.0;
  if (grad)
    for (int d = 0; d < f.dim; d++) grad[d] = 0.0;
  for (int m = 0; m < f.n_modes; m++) {
    double arg = f.phi[m];
    for (int d = 0; d < f.dim; d++) arg += 2.0 * M_PI * f.k[m * f.dim + d] * x[d];
    double sn, cs;
    sincos(arg, &sn, &cs);
    g += f.w[m] * sn;
    if (grad) {
      const double c = f.w[m] * cs;
      for (int d = 0; d < f.dim; d++) grad[d] += c * 2.0 * M_PI * f.k[m * f.dim + d];
    }
  }
  if (grad)
    for (int d = 0; d < f.dim; d++) grad[d] *= f.scale;
  return f.scale * g;
}

// The element kernels below follow the reference's arithmetic step by step
// (the plugins of element_geom.hpp, Jxw = w * jacobian * volume, K(j,k) +=
// Jxw * <g_j, g_k>) with floating-point contraction off, so each element
// matrix is bit-identical to the host restatement.
#pragma clang fp contract(off)

// ---------------------------------------------------------------------------
// element matrices: StiffMatrix / MassMatrix / variable-coefficient stiffness
// (BilinearOperator.templates.h:342-406, EllipticEquation.cpp:65-108):
//   K(j,k) += w_l * jacobian * volume [* a(x_l)] * <g_j, g_k>  (or * v_j v_k)
// one thread per element.  The matrices are stored entry-major (entry t of
// element e at Ke[t * n_elem + e]), so a wavefront's stores are contiguous.
// The stiffness forms store the upper triangle only (t over j <= k, row by
// row): <g_j,g_k> and <g_k,g_j> are the same products, so the lower triangle
// is the upper one bit for bit.  The mass form (Jxw v_j) v_k is not
// symmetric in rounding and stores all NL*NL entries.  qb holds the reference
// barycentric coordinates of the rule's points (ref_bary, computed once per
// rule), qd the points and weights.
// ---------------------------------------------------------------------------
template <int NL> __host__ __device__ constexpr int upper_index(int j, int k) { // j <= k
  return j * NL - j * (j - 1) / 2 + (k - j);
}
template <int NL> constexpr int upper_count() { return NL * (NL + 1) / 2; }

// The launch covers the elements [e0, e1) (a chunk of a pipelined assembly);
// n_elem is the stride of the entry-major store.  EM stores element-major
// instead (an element's entries contiguous: the upper triangle, or the full
// NL x NL row-major for the mass form), which suits meshes whose rows' elements
// lie far apart in index (tetrahedra), where a wavefront's loads of one entry
// over its rows' elements would each touch a separate line.
template <int KIND, int OP, bool EM>
__global__ __launch_bounds__(256) void k_element_matrices(
    int64_t n_elem, int64_t e0, int64_t e1, const double *__restrict__ coords, const int32_t *__restrict__ ev,
    const double *__restrict__ qd, const double *__restrict__ qb, int nq, FieldDev fld,
    double *__restrict__ Ke) {
  constexpr int D = KindInfo<KIND>::D, NL = KindInfo<KIND>::NL;
  const int64_t e = e0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= e1) return;
  double v[D + 1][D];
  load_elem<KIND>(e, coords, ev, v);
  const double volume = tmpl_volume<D>();
  const double jac = l2g_jacobian<D>(v);
  double K[NL][NL];
#pragma unroll
  for (int j = 0; j < NL; j++)
#pragma unroll
    for (int k = 0; k < NL; k++) K[j][k] = 0.0;
  for (int l = 0; l < nq; l++) {
    const double *q = qd + l * (D + 1);
    double x[D];
    bary_to_global<D>(qb + l * (D + 1), v, x);
    const double Jxw = q[D] * jac * volume;
    if constexpr (OP == 1) { // mass
      double val[NL];
      basis_val<KIND>(x, v, val);
#pragma unroll
      for (int j = 0; j < NL; j++)
#pragma unroll
        for (int k = 0; k < NL; k++) K[j][k] += Jxw * val[j] * val[k];
    } else {
      double g[NL][D];
      basis_grad<KIND>(x, v, g);
      double coef = 1.0;
      if constexpr (OP == 2) coef = exp(field_g(fld, x, nullptr));
#pragma unroll
      for (int j = 0; j < NL; j++)
#pragma unroll
        for (int k = j; k < NL; k++) {
          // innerProduct sums from +0.0; starting from the first product
          // instead changes ip only when that sum is +0.0 and the product
          // -0.0, and a +-0.0 term then adds to K, which starts at +0.0 and
          // never holds -0.0 (a round-to-nearest sum is -0.0 only when both
          // addends are), so every K bit is the same, with 21 fewer FP64 adds
          // per quadrature point on P2
          double ip = g[j][0] * g[k][0];
#pragma unroll
          for (int d = 1; d < D; d++) ip += g[j][d] * g[k][d];
          if constexpr (OP == 2)
            K[j][k] += Jxw * coef * ip;
          else
            K[j][k] += Jxw * ip;
        }
    }
  }
  if constexpr (OP == 1) {
#pragma unroll
    for (int j = 0; j < NL; j++)
#pragma unroll
      for (int k = 0; k < NL; k++) {
        if constexpr (EM)
          Ke[e * (NL * NL) + j * NL + k] = K[j][k];
        else
          Ke[(int64_t)(j * NL + k) * n_elem + e] = K[j][k];
      }
  } else {
#pragma unroll
    for (int j = 0; j < NL; j++)
#pragma unroll
      for (int k = j; k < NL; k++) {
        if constexpr (EM)
          Ke[e * upper_count<NL>() + upper_index<NL>(j, k)] = K[j][k];
        else
          Ke[(int64_t)upper_index<NL>(j, k) * n_elem + e] = K[j][k];
      }
  }
}

// ---------------------------------------------------------------------------
// element -> CSR gather: one thread per row, 256 rows per block; the block's
// contiguous slice of the value array is accumulated in LDS (each slot in
// ascending element order, the summation order of buildSparseMatrix) and
// written back coalesced.  The launch covers the row blocks [blk0, blk1).  NL is the element DOF count; LAYOUT names how Ke
// holds the element matrices: 0 element-major NL x NL row-major (matrices
// given by the caller, or the element-major mass form), 1 entry-major upper
// triangle (stiffness forms), 2 entry-major NL x NL (mass form), 3
// element-major upper triangle (stiffness forms).
// ---------------------------------------------------------------------------
template <int NL, int LAYOUT>
__device__ __forceinline__ double ke_entry(const double *__restrict__ Ke, int64_t n_elem, int64_t e, int a, int b) {
  if constexpr (LAYOUT == 0) return Ke[(e * NL + a) * NL + b];
  if constexpr (LAYOUT == 2) return Ke[(int64_t)(a * NL + b) * n_elem + e];
  const int j = a < b ? a : b, k = a < b ? b : a;
  if constexpr (LAYOUT == 3) return Ke[e * upper_count<NL>() + upper_index<NL>(j, k)];
  return Ke[(int64_t)upper_index<NL>(j, k) * n_elem + e];
}

// BATCH: an element's NL DOFs are distinct, so its NL slots in row i are
// distinct and the row's NL LDS updates of one element are independent: the
// NL old values are read together, then the NL sums written (one LDS round
// trip per element instead of NL dependent ones).  Each slot still takes its
// contributions in ascending element order.
constexpr int GATHER_GROUP = 4;
template <int NL, int LAYOUT, bool BATCH>
__global__ __launch_bounds__(256) void k_gather(
    int64_t n_rows, int64_t blk0, int64_t blk1, int xcd, const int64_t *__restrict__ rp,
    const int64_t *__restrict__ adj_ptr, const int32_t *__restrict__ adj, const uint8_t *__restrict__ slot,
    int64_t n_elem, const double *__restrict__ Ke, double *__restrict__ val) {
  extern __shared__ double seg[];
  // with xcd = 1 each XCD takes a contiguous range of row blocks: the rows of
  // one element (numbered close together) then read its matrix through one L2
  const int64_t blk = blk0 + logical_block(xcd);
  if (blk >= blk1) return; // the whole block (the grid is rounded up to a multiple of 8)
  const int64_t r0 = blk * 256;
  const int64_t r1 = r0 + 256 < n_rows ? r0 + 256 : n_rows;
  const int64_t base = rp[r0];
  const int64_t segn = rp[r1] - base;
  for (int64_t k = threadIdx.x; k < segn; k += 256) seg[k] = 0.0;
  __syncthreads();
  const int64_t i = r0 + threadIdx.x;
  if (i < r1) {
    double *row = seg + (rp[i] - base);
    const int64_t p0 = adj_ptr[i], p1 = adj_ptr[i + 1];
    // the row's elements in groups of GATHER_GROUP: a group's adjacency
    // entries, then its element-matrix rows and slots, are loaded together
    // (two dependent loads per group instead of per element); the group's
    // elements are then summed one after the other in ascending order
    for (int64_t pg = p0; pg < p1; pg += GATHER_GROUP) {
      int32_t ea[GATHER_GROUP];
      int s[GATHER_GROUP][NL];
      double kv[GATHER_GROUP][NL];
#pragma unroll
      for (int g = 0; g < GATHER_GROUP; g++) ea[g] = pg + g < p1 ? adj[pg + g] : 0;
#pragma unroll
      for (int g = 0; g < GATHER_GROUP; g++)
        if (pg + g < p1) {
          const int64_t e = ea[g] >> 3;
          const int a = ea[g] & 7;
          const uint8_t *sl = slot + (pg + g) * NL;
#pragma unroll
          for (int b = 0; b < NL; b++) {
            s[g][b] = sl[b];
            kv[g][b] = ke_entry<NL, LAYOUT>(Ke, n_elem, e, a, b);
          }
        }
#pragma unroll
      for (int g = 0; g < GATHER_GROUP; g++)
        if (pg + g < p1) {
          if constexpr (BATCH) {
            double old[NL];
#pragma unroll
            for (int b = 0; b < NL; b++) old[b] = row[s[g][b]];
#pragma unroll
            for (int b = 0; b < NL; b++) row[s[g][b]] = old[b] + kv[g][b];
          } else {
#pragma unroll
            for (int b = 0; b < NL; b++) row[s[g][b]] += kv[g][b];
          }
        }
    }
  }
  __syncthreads();
  for (int64_t k = threadIdx.x; k < segn; k += 256) val[base + k] = seg[k];
}

// The same gather with the block's element contributions staged in LDS:
// the block's rows own the contiguous adjacency range [adj_ptr[r0],
// adj_ptr[r1]), which is walked in chunks of GATHER_CHUNK entries.  All 256
// threads load a chunk's element-matrix rows and slots together (each thread
// several independent entries, so many loads are in flight), then each row
// sums its own entries of the chunk from LDS in ascending order.  Chunks go
// in order, so every slot still takes its contributions in ascending element
// order.  For rows with long element lists (tetrahedra, ~24), where the
// one-thread walk is a chain of dependent loads.  LDS: seg_cap doubles of
// row values, then the chunk's NL values and NL slots per entry; the slots of
// one element of a row are distinct (BATCH's condition).
constexpr int GATHER_CHUNK = 1024;
template <int NL, int LAYOUT>
__global__ __launch_bounds__(256) void k_gather_staged(
    int64_t n_rows, int64_t blk0, int64_t blk1, int xcd, int64_t seg_cap, const int64_t *__restrict__ rp,
    const int64_t *__restrict__ adj_ptr, const int32_t *__restrict__ adj, const uint8_t *__restrict__ slot,
    int64_t n_elem, const double *__restrict__ Ke, double *__restrict__ val) {
  extern __shared__ double seg[];
  double *skv = seg + seg_cap;
  uint8_t *ssl = reinterpret_cast<uint8_t *>(skv + GATHER_CHUNK * NL);
  const int64_t blk = blk0 + logical_block(xcd);
  if (blk >= blk1) return; // the whole block (the grid is rounded up to a multiple of 8)
  const int64_t r0 = blk * 256;
  const int64_t r1 = r0 + 256 < n_rows ? r0 + 256 : n_rows;
  const int64_t base = rp[r0];
  const int64_t segn = rp[r1] - base;
  for (int64_t k = threadIdx.x; k < segn; k += 256) seg[k] = 0.0;
  const int64_t i = r0 + threadIdx.x;
  const bool mine = i < r1;
  double *row = mine ? seg + (rp[i] - base) : seg;
  const int64_t pa = mine ? adj_ptr[i] : 0, pb = mine ? adj_ptr[i + 1] : 0;
  const int64_t q0 = adj_ptr[r0], q1 = adj_ptr[r1];
  for (int64_t c0 = q0; c0 < q1; c0 += GATHER_CHUNK) {
    const int64_t c1 = c0 + GATHER_CHUNK < q1 ? c0 + GATHER_CHUNK : q1;
    __syncthreads(); // the previous chunk is summed (and the row values cleared)
    for (int64_t p = c0 + threadIdx.x; p < c1; p += 256) {
      const int32_t ea = adj[p];
      const int64_t e = ea >> 3;
      const int a = ea & 7;
      const int k = (int)(p - c0);
#pragma unroll
      for (int b = 0; b < NL; b++) {
        skv[k * NL + b] = ke_entry<NL, LAYOUT>(Ke, n_elem, e, a, b);
        ssl[k * NL + b] = slot[p * NL + b];
      }
    }
    __syncthreads();
    const int64_t lo = pa > c0 ? pa : c0, hi = pb < c1 ? pb : c1;
    for (int64_t p = lo; p < hi; p++) {
      const int k = (int)(p - c0);
      int sl[NL];
      double old[NL];
#pragma unroll
      for (int b = 0; b < NL; b++) sl[b] = ssl[k * NL + b];
#pragma unroll
      for (int b = 0; b < NL; b++) old[b] = row[sl[b]];
#pragma unroll
      for (int b = 0; b < NL; b++) row[sl[b]] = old[b] + skv[k * NL + b];
    }
  }
  __syncthreads();
  for (int64_t k = threadIdx.x; k < segn; k += 256) val[base + k] = seg[k];
}

// ---------------------------------------------------------------------------
// two FE spaces on one mesh (BilinearOperator's same-mesh branch):
// L2InnerProduct's element matrix (BilinearOperator.templates.h:300-326, the
// quadrature and Jacobian of element 0): K(j,k) += Jxw * v0_j * v1_k,
// stored entry-major (entry j * N1 + k of element e at Ke[.. * n_elem + e])
// ---------------------------------------------------------------------------
template <int K0, int K1>
__global__ __launch_bounds__(256) void k_element_mass_pair(int64_t n_elem, const double *__restrict__ coords,
                                                           const int32_t *__restrict__ ev,
                                                           const double *__restrict__ qd,
                                                           const double *__restrict__ qb, int nq,
                                                           double *__restrict__ Ke) {
  constexpr int D = KindInfo<K0>::D, N0 = KindInfo<K0>::NL, N1 = KindInfo<K1>::NL;
  static_assert(KindInfo<K1>::D == D, "the two spaces share the mesh");
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_elem) return;
  double v[D + 1][D];
  load_elem<K0>(e, coords, ev, v);
  const double volume = tmpl_volume<D>();
  const double jac = l2g_jacobian<D>(v);
  double K[N0][N1];
#pragma unroll
  for (int j = 0; j < N0; j++)
#pragma unroll
    for (int k = 0; k < N1; k++) K[j][k] = 0.0;
  for (int l = 0; l < nq; l++) {
    double x[D];
    bary_to_global<D>(qb + l * (D + 1), v, x);
    const double Jxw = qd[l * (D + 1) + D] * jac * volume;
    double v0[N0], v1[N1];
    basis_val<K0>(x, v, v0);
    basis_val<K1>(x, v, v1);
#pragma unroll
    for (int j = 0; j < N0; j++)
#pragma unroll
      for (int k = 0; k < N1; k++) K[j][k] += Jxw * v0[j] * v1[k];
  }
#pragma unroll
  for (int j = 0; j < N0; j++)
#pragma unroll
    for (int k = 0; k < N1; k++) Ke[(int64_t)(j * N1 + k) * n_elem + e] = K[j][k];
}

// the rectangular gather: row i of space 0 sums, in ascending element
// order, row a of each adjacent element matrix (N1 entries) into its slots.
// LAYOUT 0: element-major N0 x N1 row-major (caller's matrices), 2:
// entry-major as k_element_mass_pair writes them.
template <int N0, int N1, int LAYOUT>
__global__ __launch_bounds__(256) void k_gather_pair(int64_t n_rows, const int64_t *__restrict__ rp,
                                                     const int64_t *__restrict__ adj_ptr,
                                                     const int32_t *__restrict__ adj,
                                                     const uint8_t *__restrict__ slot, int64_t n_elem,
                                                     const double *__restrict__ Ke, double *__restrict__ val) {
  extern __shared__ double seg[];
  const int64_t r0 = (int64_t)blockIdx.x * 256;
  const int64_t r1 = r0 + 256 < n_rows ? r0 + 256 : n_rows;
  const int64_t base = rp[r0];
  const int64_t segn = rp[r1] - base;
  for (int64_t k = threadIdx.x; k < segn; k += 256) seg[k] = 0.0;
  __syncthreads();
  const int64_t i = r0 + threadIdx.x;
  if (i < r1) {
    double *row = seg + (rp[i] - base);
    for (int64_t p = adj_ptr[i]; p < adj_ptr[i + 1]; p++) {
      const int32_t ea = adj[p];
      const int64_t e = ea >> 3;
      const int a = ea & 7;
      const uint8_t *sl = slot + p * N1;
      double kv[N1];
#pragma unroll
      for (int b = 0; b < N1; b++)
        kv[b] = LAYOUT == 0 ? Ke[(e * N0 + a) * N1 + b] : Ke[(int64_t)(a * N1 + b) * n_elem + e];
#pragma unroll
      for (int b = 0; b < N1; b++) row[sl[b]] += kv[b];
    }
  }
  __syncthreads();
  for (int64_t k = threadIdx.x; k < segn; k += 256) val[base + k] = seg[k];
}

// ---------------------------------------------------------------------------
// right-hand side: element load vectors, then a per-DOF gather in ascending
// element order (Operator::L2Discretize, Operator.discretize.templates.h:80-152)
// ---------------------------------------------------------------------------
template <int KIND>
__global__ __launch_bounds__(256) void k_element_loads(
    int64_t n_elem, const double *__restrict__ coords, const int32_t *__restrict__ ev,
    const int32_t *__restrict__ edof, const double *__restrict__ qd, int nq, int fkind,
    FieldDev fld, const double *__restrict__ nodal, double *__restrict__ Fe) {
  constexpr int D = KindInfo<KIND>::D, NL = KindInfo<KIND>::NL;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_elem) return;
  double v[D + 1][D];
  load_elem<KIND>(e, coords, ev, v);
  const double volume = tmpl_volume<D>();
  const double jac = l2g_jacobian<D>(v);
  double F[NL];
#pragma unroll
  for (int j = 0; j < NL; j++) F[j] = 0.0;
  for (int l = 0; l < nq; l++) {
    const double *q = qd + l * (D + 1);
    double x[D];
    local_to_global<D>(q, v, x);
    double val[NL];
    basis_val<KIND>(x, v, val);
    double f;
    if (fkind == 3) {
      f = 1.0;
    } else if (fkind == 2) {
      f = 0.0;
#pragma unroll
      for (int j = 0; j < NL; j++) f += nodal[edof[e * NL + j]] * val[j];
    } else {
      double u = 1.0, gu[D];
#pragma unroll
      for (int d = 0; d < D; d++) u *= sin(M_PI * x[d]);
      if (fkind == 0) {
        f = D * M_PI * M_PI * u;
      } else {
#pragma unroll
        for (int d = 0; d < D; d++) {
          double gd = M_PI * cos(M_PI * x[d]);
#pragma unroll
          for (int c = 0; c < D; c++)
            if (c != d) gd *= sin(M_PI * x[c]);
          gu[d] = gd;
        }
        double gg[3];
        const double g = field_g(fld, x, gg);
        double dot = 0;
#pragma unroll
        for (int d = 0; d < D; d++) dot += gg[d] * gu[d];
        f = exp(g) * (D * M_PI * M_PI * u - dot);
      }
    }
    const double Jxw = q[D] * jac * volume;
#pragma unroll
    for (int j = 0; j < NL; j++) F[j] += Jxw * f * val[j];
  }
#pragma unroll
  for (int j = 0; j < NL; j++) Fe[e * NL + j] = F[j];
}

// load vectors from f given at the physical quadrature points, fq[e*nq + l]
// (Operator::L2Discretize with a host function, Operator.discretize.templates.h:121-152)
template <int KIND>
__global__ __launch_bounds__(256) void k_element_loads_qp(int64_t n_elem, const double *__restrict__ coords,
                                                          const int32_t *__restrict__ ev,
                                                          const double *__restrict__ qd, int nq,
                                                          const double *__restrict__ fq, double *__restrict__ Fe) {
  constexpr int D = KindInfo<KIND>::D, NL = KindInfo<KIND>::NL;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_elem) return;
  double v[D + 1][D];
  load_elem<KIND>(e, coords, ev, v);
  const double volume = tmpl_volume<D>();
  const double jac = l2g_jacobian<D>(v);
  double F[NL];
#pragma unroll
  for (int j = 0; j < NL; j++) F[j] = 0.0;
  for (int l = 0; l < nq; l++) {
    const double *q = qd + l * (D + 1);
    double x[D];
    local_to_global<D>(q, v, x);
    double val[NL];
    basis_val<KIND>(x, v, val);
    const double f = fq[e * nq + l];
    const double Jxw = q[D] * jac * volume;
#pragma unroll
    for (int j = 0; j < NL; j++) F[j] += Jxw * f * val[j];
  }
#pragma unroll
  for (int j = 0; j < NL; j++) Fe[e * NL + j] = F[j];
}

// physical quadrature points and Jxw of every element (the host evaluates a
// user function there)
template <int KIND>
__global__ __launch_bounds__(256) void k_qpoints(int64_t n_elem, const double *__restrict__ coords,
                                                 const int32_t *__restrict__ ev, const double *__restrict__ qd,
                                                 int nq, double *__restrict__ pts, double *__restrict__ jxw) {
  constexpr int D = KindInfo<KIND>::D;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_elem) return;
  double v[D + 1][D];
  load_elem<KIND>(e, coords, ev, v);
  const double volume = tmpl_volume<D>();
  const double jac = l2g_jacobian<D>(v);
  for (int l = 0; l < nq; l++) {
    const double *q = qd + l * (D + 1);
    double x[D];
    local_to_global<D>(q, v, x);
#pragma unroll
    for (int d = 0; d < D; d++) pts[(e * nq + l) * D + d] = x[d];
    if (jxw) jxw[e * nq + l] = q[D] * jac * volume;
  }
}

// errors against an exact solution given at the quadrature points:
// h1 = 0: uq[e*nq+l]; h1 = 1: gq[(e*nq+l)*D + d]
template <int KIND>
__device__ __forceinline__ double element_error_qp(int64_t e, const double *__restrict__ coords,
                                                   const int32_t *__restrict__ ev,
                                                   const int32_t *__restrict__ edof,
                                                   const double *__restrict__ qd, int nq, int h1,
                                                   const double *__restrict__ uh,
                                                   const double *__restrict__ exact) {
  constexpr int D = KindInfo<KIND>::D, NL = KindInfo<KIND>::NL;
  double v[D + 1][D];
  load_elem<KIND>(e, coords, ev, v);
  const double volume = tmpl_volume<D>();
  const double jac = l2g_jacobian<D>(v);
  double u_loc[NL];
#pragma unroll
  for (int j = 0; j < NL; j++) u_loc[j] = uh[edof[e * NL + j]];
  double err = 0.0;
  for (int l = 0; l < nq; l++) {
    const double *q = qd + l * (D + 1);
    double x[D];
    local_to_global<D>(q, v, x);
    const double Jxw = q[D] * jac * volume;
    if (!h1) {
      double val[NL];
      basis_val<KIND>(x, v, val);
      double uhv = 0.0;
#pragma unroll
      for (int j = 0; j < NL; j++) uhv += u_loc[j] * val[j];
      const double df = exact[e * nq + l] - uhv;
      err += Jxw * df * df;
    } else {
      double g[NL][D];
      basis_grad<KIND>(x, v, g);
#pragma unroll
      for (int d = 0; d < D; d++) {
        double gh = 0.0;
#pragma unroll
        for (int j = 0; j < NL; j++) gh += u_loc[j] * g[j][d];
        const double df = gh - exact[(e * nq + l) * D + d];
        err += Jxw * df * df;
      }
    }
  }
  return err;
}

__global__ __launch_bounds__(256) void k_gather_vec(int64_t n, int nl, const int64_t *__restrict__ adj_ptr,
                                                    const int32_t *__restrict__ adj,
                                                    const double *__restrict__ Fe, double *__restrict__ f) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int64_t p = adj_ptr[i]; p < adj_ptr[i + 1]; p++) {
    const int32_t ea = adj[p];
    s += Fe[(int64_t)(ea >> 3) * nl + (ea & 7)];
  }
  f[i] = s;
}

// BoundaryConditionAdmin::apply with zero Dirichlet data
// (FEMSpace.templates.h:1338-1373): a marked row keeps only its diagonal and
// gets f_i = A_ii * 0; the eliminated column entries A_ki are zeroed (the
// update f_k -= A_ki * 0 / A_ii leaves f_k unchanged).
__global__ __launch_bounds__(256) void k_dirichlet_zero(int64_t n, const int64_t *__restrict__ rp,
                                                        const int32_t *__restrict__ col,
                                                        const int32_t *__restrict__ mark,
                                                        double *__restrict__ val, double *__restrict__ f) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (mark[i] != 0) {
    f[i] = val[rp[i]] * 0.0;
    for (int64_t j = rp[i] + 1; j < rp[i + 1]; j++) val[j] = 0.0;
  } else {
    for (int64_t j = rp[i] + 1; j < rp[i + 1]; j++)
      if (mark[col[j]] != 0) val[j] = 0.0;
  }
}

// BoundaryConditionAdmin::apply with Dirichlet data g on the flagged DOFs
// (FEMSpace.templates.h:1338-1373): a flagged row gets f_i = A_ii g_i and
// keeps only its diagonal; with sym (preserve_symmetry) every other row k
// subtracts, column by column in ascending order (the order in which the
// reference visits the boundary DOFs), A_ki * (A_ii g_i) / A_ii and zeroes
// A_ki, without it the other rows stay as they are.  Flagged rows read the
// pre-elimination values of their own row only.
__global__ __launch_bounds__(256) void k_dirichlet_values(int64_t n, const int64_t *__restrict__ rp,
                                                          const int32_t *__restrict__ col,
                                                          const uint8_t *__restrict__ flag,
                                                          const double *__restrict__ diag0,
                                                          const double *__restrict__ g,
                                                          double *__restrict__ val, double *__restrict__ f, int sym) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (flag[i]) {
    f[i] = val[rp[i]] * g[i];
    for (int64_t j = rp[i] + 1; j < rp[i + 1]; j++) val[j] = 0.0;
  } else if (sym) {
    double fi = f[i];
    for (int64_t j = rp[i] + 1; j < rp[i + 1]; j++) {
      const int32_t c = col[j];
      if (flag[c]) {
        const double fc = diag0[c] * g[c];
        fi -= val[j] * fc / diag0[c];
        val[j] = 0.0;
      }
    }
    f[i] = fi;
  }
}

// diagonal of a CSR matrix whose rows start with the diagonal
__global__ __launch_bounds__(256) void k_csr_diag(int64_t n, const int64_t *__restrict__ rp,
                                                  const double *__restrict__ val, double *__restrict__ d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = val[rp[i]];
}

// ---------------------------------------------------------------------------
// error functionals (Functional.templates.h:296-317,399-426): per-element
// contributions, then a fixed-order reduction
// ---------------------------------------------------------------------------
template <int KIND>
__global__ __launch_bounds__(256) void k_element_error(
    int64_t n_elem, const double *__restrict__ coords, const int32_t *__restrict__ ev,
    const int32_t *__restrict__ edof, const double *__restrict__ qd, int nq, int h1,
    const double *__restrict__ uh, double *__restrict__ out) {
  constexpr int D = KindInfo<KIND>::D, NL = KindInfo<KIND>::NL;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_elem) return;
  double v[D + 1][D];
  load_elem<KIND>(e, coords, ev, v);
  const double volume = tmpl_volume<D>();
  const double jac = l2g_jacobian<D>(v);
  double u_loc[NL];
#pragma unroll
  for (int j = 0; j < NL; j++) u_loc[j] = uh[edof[e * NL + j]];
  double err = 0.0;
  for (int l = 0; l < nq; l++) {
    const double *q = qd + l * (D + 1);
    double x[D];
    local_to_global<D>(q, v, x);
    const double Jxw = q[D] * jac * volume;
    if (!h1) {
      double val[NL];
      basis_val<KIND>(x, v, val);
      double uhv = 0.0, ue = 1.0;
#pragma unroll
      for (int j = 0; j < NL; j++) uhv += u_loc[j] * val[j];
#pragma unroll
      for (int d = 0; d < D; d++) ue *= sin(M_PI * x[d]);
      err += Jxw * (ue - uhv) * (ue - uhv);
    } else {
      double g[NL][D];
      basis_grad<KIND>(x, v, g);
#pragma unroll
      for (int d = 0; d < D; d++) {
        double gh = 0.0;
#pragma unroll
        for (int j = 0; j < NL; j++) gh += u_loc[j] * g[j][d];
        double ge = M_PI * cos(M_PI * x[d]);
#pragma unroll
        for (int c = 0; c < D; c++)
          if (c != d) ge *= sin(M_PI * x[c]);
        err += Jxw * (gh - ge) * (gh - ge);
      }
    }
  }
  out[e] = err;
}

#pragma clang fp contract(on)

// ---------------------------------------------------------------------------
// fixed-order reductions
// ---------------------------------------------------------------------------
__device__ inline double block_sum(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// partial sums of x over a fixed grid (grid-stride), one per block
__global__ __launch_bounds__(256) void k_sum_partial(int64_t n, const double *__restrict__ x,
                                                     double *__restrict__ part) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    s += x[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// out[k] = sum of part[k*stride + 0 .. np) for k < nsum, one block
// one workgroup of FIN_THREADS sums np block partials per scalar (fixed
// order: strided per-thread sums, then a tree)
constexpr int FIN_THREADS = 1024;

// a thread's strided sum of part[i], i = t, t + stride, ... < np, in that
// order; the loads are issued eight at a time
__device__ inline double strided_sum(const double *__restrict__ part, int np, int t, int stride) {
  double s = 0.0;
  for (int i0 = t; i0 < np; i0 += 8 * stride) {
    double v[8];
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (i0 + k * stride < np) v[k] = part[i0 + k * stride];
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (i0 + k * stride < np) s += v[k];
  }
  return s;
}
__global__ __launch_bounds__(FIN_THREADS) void k_finalize(const double *__restrict__ part, int np, int nsum,
                                                          double *__restrict__ out) {
  __shared__ double sh[FIN_THREADS];
  for (int k = 0; k < nsum; k++) {
    double s = strided_sum(part + (int64_t)k * np, np, threadIdx.x, blockDim.x);
    s = block_sum(s, sh);
    if (threadIdx.x == 0) out[k] = s;
  }
}

// Functional::L2Error / H1SemiError (Functional.templates.h:296-317,399-426)
// with the exact values at the quadrature points given: each thread sums the
// squared errors of its elements (grid-stride, ascending), the block adds
// its threads' sums in a fixed tree, and part[block] receives the block's
// sum; gridDim.x partial sums, no per-element array
template <int KIND>
__global__ __launch_bounds__(256) void k_error_qp_part(int64_t n_elem, const double *__restrict__ coords,
                                                       const int32_t *__restrict__ ev,
                                                       const int32_t *__restrict__ edof,
                                                       const double *__restrict__ qd, int nq, int h1,
                                                       const double *__restrict__ uh,
                                                       const double *__restrict__ exact,
                                                       double *__restrict__ part) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n_elem;
       e += (int64_t)gridDim.x * blockDim.x)
    s += element_error_qp<KIND>(e, coords, ev, edof, qd, nq, h1, uh, exact);
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// out = sqrt(|sum of part[0 .. np)|), the sum in k_finalize's order and the
// square root correctly rounded (the value a host sqrt gives)
__global__ __launch_bounds__(FIN_THREADS) void k_finalize_sqrt(const double *__restrict__ part, int np,
                                                               double *__restrict__ out) {
  __shared__ double sh[FIN_THREADS];
  double s = strided_sum(part, np, threadIdx.x, blockDim.x);
  s = block_sum(s, sh);
  if (threadIdx.x == 0) out[0] = __dsqrt_rn(fabs(s));
}

// ---------------------------------------------------------------------------
// sliced-ELL kernels (one wavefront per slice, one lane per row)
// ---------------------------------------------------------------------------
// Pointers with their address space in the type: loads and stores through
// them compile to global_* (or ds_*) instructions.  Pointers read from
// memory (tables of levels or members) are otherwise generic, and their
// accesses compile to flat instructions, which count in the LDS counter as
// well and are waited for more coarsely.
template <class T> using GPtr = __attribute__((address_space(1))) T *;
template <class T> using LPtr = __attribute__((address_space(3))) T *;
template <class T> __device__ __forceinline__ GPtr<T> gptr(T *p) { return (GPtr<T>)p; }
template <class T> __device__ __forceinline__ GPtr<const T> gptr(const T *p) { return (GPtr<const T>)p; }

// a slice descriptor read through an address-space-qualified pointer
// (field by field: the implicit copy works on generic references only)
template <class SP> __device__ __forceinline__ SliceInfo load_si(SP p, int s) {
  SliceInfo q;
  q.row0 = p[s].row0;
  q.nrow = p[s].nrow;
  q.width = p[s].width;
  q.off = p[s].off;
  return q;
}

// streamed matrix data: non-temporal loads keep the gathered vector in cache
template <class U> __device__ inline U ld_stream(const U *p) { return __builtin_nontemporal_load(p); }

template <bool NT, class P> __device__ inline auto ld_mat(P p) {
  if constexpr (NT)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}

// one row of a sliced-ELL matrix times x; NT selects non-temporal loads for
// the matrix stream (the level-0 matrices, larger than the caches), plain
// loads for the small levels whose matrices stay cached between cycles.
// The stored entries are taken in chunks of CH: all column and value loads
// of a chunk are issued before the first gather and all gathers before the
// first product, so a row costs two dependent memory latencies per chunk
// rather than two per few entries.  The products are summed one by one in
// stored order, the same sum for every chunk size.
// chunk sizes measured on the 4.2M-DOF P2 system: 16 for the V-cycle
// kernels (the small levels are latency-bound; the level-0 colour sweeps gain
// too), 8 for the natural-order FP64 SpMV family, where the registers of a
// longer chunk cost more occupancy than they save in latency
constexpr int SELL_CHUNK = 16;
constexpr int SELL_CHUNK_SPMV = 8;
// per-level launches: the level-0 streams (NT) use SELL_CHUNK; the small
// FP32 levels take a whole row in one chunk (their rows are short, and the
// launches are few waves long, so registers are free)
template <class T, bool NT> struct LaunchChunk {
  static constexpr int v = (!NT && sizeof(T) == 4) ? 24 : SELL_CHUNK;
};
// Columns are int32 indices, or (CI = uint16_t) offsets from the slice's
// base column cbase.  The pointers may carry an address space (the batch
// kernels pass global-typed pointers read from their tables).
template <class T, bool NT = false, int CH = SELL_CHUNK, class CI = int32_t, class CP = const CI *,
          class VP = const T *, class XP = const T *>
__device__ inline T sell_dot(const SliceInfo &s, int lane, CP __restrict__ cols, VP __restrict__ vals,
                             XP __restrict__ x, int32_t cbase = 0) {
  const auto c = cols + s.off + lane;
  const auto v = vals + s.off + lane;
  T acc = 0;
  for (int j = 0; j < s.width; j += CH) {
    int32_t cc[CH]; // the base is added at the gather, after every load of the chunk is issued
    T vv[CH], xx[CH];
#pragma unroll
    for (int k = 0; k < CH; k++)
      if (j + k < s.width) {
        cc[k] = ld_mat<NT>(c + (j + k) * WAVE);
        vv[k] = ld_mat<NT>(v + (j + k) * WAVE);
      }
#pragma unroll
    for (int k = 0; k < CH; k++)
      if (j + k < s.width) xx[k] = x[cbase + cc[k]];
#pragma unroll
    for (int k = 0; k < CH; k++)
      if (j + k < s.width) acc += vv[k] * xx[k];
  }
  return acc;
}

// Colour-coded 16-bit columns (the streamed level-0 colour layout): entry
// e = (colour << cshift) | offset, column = cb[colour] + offset, where cb
// holds the slice's base position in each colour's range (8 per slice).
// The neighbours of a slice's rows sit in a narrow band of every other
// colour's range, so the offsets fit in 13 bits on the P2 and tetrahedral
// systems.  Same chunked loads and the same summation order as sell_dot.
constexpr int CC_BASES = 8;
template <class T, bool NT = true, int CH = SELL_CHUNK>
__device__ inline T sell_dot_cc(const SliceInfo &s, int lane, const uint16_t *__restrict__ cols,
                                const int32_t *__restrict__ cb, int cshift, const T *__restrict__ vals,
                                const T *__restrict__ x) {
  int32_t b[CC_BASES];
#pragma unroll
  for (int k = 0; k < CC_BASES; k++) b[k] = cb[k];
  const int32_t mask = (1 << cshift) - 1;
  const uint16_t *c = cols + s.off + lane;
  const T *v = vals + s.off + lane;
  T acc = 0;
  for (int j = 0; j < s.width; j += CH) {
    int32_t cc[CH];
    T vv[CH], xx[CH];
#pragma unroll
    for (int k = 0; k < CH; k++)
      if (j + k < s.width) {
        cc[k] = ld_mat<NT>(c + (j + k) * WAVE);
        vv[k] = ld_mat<NT>(v + (j + k) * WAVE);
      }
#pragma unroll
    for (int k = 0; k < CH; k++)
      if (j + k < s.width) {
        const int ci = cc[k] >> cshift;
        int32_t base = b[0];
#pragma unroll
        for (int m = 1; m < CC_BASES; m++) base = ci == m ? b[m] : base;
        xx[k] = x[base + (cc[k] & mask)];
      }
#pragma unroll
    for (int k = 0; k < CH; k++)
      if (j + k < s.width) acc += vv[k] * xx[k];
  }
  return acc;
}

__device__ inline int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE); }

// DPP lane moves inside rows of 16 lanes (VALU, no LDS round trip):
// quad_perm [1,0,3,2] and [2,3,0,1], row_half_mirror, row_mirror
constexpr int DPP_QUAD_XOR1 = 0xB1, DPP_QUAD_XOR2 = 0x4E, DPP_HALF_MIRROR = 0x141, DPP_MIRROR = 0x140;
template <int CTRL> __device__ inline float dpp_move(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL> __device__ inline double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp((int)b, (int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(b >> 32), (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// A row of a slice with 2^gs lanes per row: lane = row_in_slice * 2^gs + k,
// and sell_dot of lane gives the partial sum of the row's entries k, k + 2^gs,
// ... in stored order.  The row's sum adds the partials pairwise over lane
// distance 1, 2, 4, 8 (an xor butterfly): lane i adds the value of lane
// i ^ 1, then of i ^ 2 (quad permutations); at distance 4 and 8 every quad
// (then every half row) already holds one value, so the mirror moves read
// the same bits as lane i ^ 4 (i ^ 8) would hold.  a + b and b + a are the
// same bits, so every lane of the group ends with the same sum.  gs = 0 is
// the one-lane-per-row sum unchanged.  The groups of a wavefront are whole
// (a slice's rows are aligned groups), so every lane read is active.
template <class T> __device__ inline T row_group_sum(T v, int gs) {
  if (gs >= 1) v += dpp_move<DPP_QUAD_XOR1>(v);
  if (gs >= 2) v += dpp_move<DPP_QUAD_XOR2>(v);
  if (gs >= 3) v += dpp_move<DPP_HALF_MIRROR>(v);
  if (gs >= 4) v += dpp_move<DPP_MIRROR>(v);
  return v;
}

// slice schedule of the grid-stride level-0 kernels: sched 0 = contiguous
// chunk per block, 1 = contiguous chunk per XCD, 2 = interleaved (the blocks
// running at one time read neighbouring slices)
__device__ inline void slice_range(int ns, int sched, int &s, int &end, int &step) {
  if (sched == 2) {
    s = blockIdx.x * WPB + wave_id();
    end = ns;
    step = gridDim.x * WPB;
  } else {
    const int spb = (ns + gridDim.x - 1) / gridDim.x;
    const int lb = logical_block(sched);
    s = lb * spb + wave_id();
    end = min(ns, (lb + 1) * spb);
    step = WPB;
  }
}

// Gauss-Seidel over the slices [s0, s1) of one colour:
// x_r = (b_r - sum_{j != r} a_rj x_j) / a_rr
// (c16 != null: colour-coded 16-bit columns c16 / cb16 / cshift, see
// sell_dot_cc; otherwise the int32 columns; 2^gs lanes per row).  No row of
// the colour reads another's x, so the group's leader writes x_r once every
// lane of the group has read (the butterfly orders it after those reads).
template <class T, bool NT>
__global__ __launch_bounds__(BLK) void k_gs(const SliceInfo *__restrict__ si, int s0, int s1,
                                            const int32_t *__restrict__ cols, const uint16_t *__restrict__ c16,
                                            const int32_t *__restrict__ cb16, int cshift,
                                            const T *__restrict__ vals, const T *__restrict__ invd,
                                            const T *__restrict__ b, T *__restrict__ x, int gs, int xcd) {
  const int s = s0 + logical_block(xcd) * WPB + wave_id();
  if (s >= s1) return;
  const SliceInfo q = si[s];
  const int lane = threadIdx.x % WAVE;
  if ((lane >> gs) >= q.nrow) return;
  const int r = q.row0 + (lane >> gs);
  const T br = b[r], ir = invd[r]; // issued before the row's dependent loads
  T acc = c16 ? sell_dot_cc<T, NT, LaunchChunk<T, NT>::v>(q, lane, c16, cb16 + (int64_t)s * CC_BASES, cshift,
                                                          vals, x)
              : sell_dot<T, NT, LaunchChunk<T, NT>::v>(q, lane, cols, vals, x);
  acc = row_group_sum(acc, gs);
  if ((lane & ((1 << gs) - 1)) == 0) x[r] = (br - acc) * ir;
}

// the first colour of a sweep that starts from x = 0 (level 0 of the
// preconditioner): every neighbour of a colour-0 row is still zero, so the
// row takes (b - 0) * invd = b * invd, bit for bit what k_gs computes; the
// rows of the other colours get the zero the sweep starts from.  Colour-0
// rows come first in the colour order (rows [0, n0)).
template <class T>
__global__ __launch_bounds__(256) void k_gs_first(int64_t n, int64_t n0, const T *__restrict__ b,
                                                  const T *__restrict__ invd, T *__restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = i < n0 ? b[i] * invd[i] : (T)0;
}

// t = b - A x (A = diag + stored off-diagonals).  Slices from s_zero on
// hold the rows of the colour a forward sweep updated last: none of their
// neighbours changed after them, so their residual is exactly zero in exact
// arithmetic and t = 0 is written without reading the matrix (s_zero = ns
// computes every row).
template <class T, bool NT>
__global__ __launch_bounds__(BLK) void k_residual(const SliceInfo *__restrict__ si, int ns, int s_zero,
                                                  const int32_t *__restrict__ cols, const uint16_t *__restrict__ c16,
                                                  const int32_t *__restrict__ cb16, int cshift,
                                                  const T *__restrict__ vals, const T *__restrict__ diag,
                                                  const T *__restrict__ b, const T *__restrict__ x,
                                                  T *__restrict__ t, int gs, int xcd) {
  const int s = logical_block(xcd) * WPB + wave_id();
  if (s >= ns) return;
  const SliceInfo q = si[s];
  const int lane = threadIdx.x % WAVE;
  if ((lane >> gs) >= q.nrow) return;
  const int r = q.row0 + (lane >> gs);
  if (s >= s_zero) {
    if ((lane & ((1 << gs) - 1)) == 0) t[r] = 0;
    return;
  }
  const T br = b[r], dr = diag[r], xr = x[r];
  T acc = c16 ? sell_dot_cc<T, NT, LaunchChunk<T, NT>::v>(q, lane, c16, cb16 + (int64_t)s * CC_BASES, cshift,
                                                          vals, x)
              : sell_dot<T, NT, LaunchChunk<T, NT>::v>(q, lane, cols, vals, x);
  acc = row_group_sum(acc, gs);
  if ((lane & ((1 << gs) - 1)) == 0) t[r] = br - dr * xr - acc;
}

// restriction: bc = P t, and the coarse guess xc = 0.  With invd_c (the
// coarse level is smoothed next, from zero), the first colour's rows
// [0, rows_c0) get xc = bc / diag at once: their neighbours are all zero, so
// this is the first colour step of the coarse sweep (as k_gs_first), and
// that launch is left out.
template <class T, bool NT>
__global__ __launch_bounds__(BLK) void k_restrict(const SliceInfo *__restrict__ si, int ns,
                                                  const int32_t *__restrict__ cols, const T *__restrict__ vals,
                                                  const T *__restrict__ t, T *__restrict__ bc,
                                                  T *__restrict__ xc, const T *__restrict__ invd_c, int rows_c0,
                                                  int gs, int xcd) {
  const int s = logical_block(xcd) * WPB + wave_id();
  if (s >= ns) return;
  const SliceInfo q = si[s];
  const int lane = threadIdx.x % WAVE;
  if ((lane >> gs) >= q.nrow) return;
  const int r = q.row0 + (lane >> gs);
  const T acc = row_group_sum(sell_dot<T, NT, LaunchChunk<T, NT>::v>(q, lane, cols, vals, t), gs);
  if ((lane & ((1 << gs) - 1)) == 0) {
    bc[r] = acc;
    xc[r] = invd_c && r < rows_c0 ? acc * invd_c[r] : (T)0;
  }
}

// prolongation: x += P^T xc (explicit transpose, no atomics).  Rows in
// [skip0, skip1) are left alone: they form the colour the post-smoother
// updates first, which overwrites them without reading them (a row's own x
// is not among its stored off-diagonal entries).
template <class T, bool NT>
__global__ __launch_bounds__(BLK) void k_prolong(const SliceInfo *__restrict__ si, int ns,
                                                 const int32_t *__restrict__ cols, const T *__restrict__ vals,
                                                 const T *__restrict__ xc, T *__restrict__ x, int skip0, int skip1,
                                                 int xcd) {
  const int s = logical_block(xcd) * WPB + wave_id();
  if (s >= ns) return;
  const SliceInfo q = si[s];
  const int lane = threadIdx.x % WAVE;
  if (lane >= q.nrow) return;
  const int r = q.row0 + lane;
  if (r >= skip0 && r < skip1) return;
  const T xr = x[r];
  x[r] = xr + sell_dot<T, NT, LaunchChunk<T, NT>::v>(q, lane, cols, vals, xc);
}

// ---------------------------------------------------------------------------
// single-workgroup V-cycle over the small levels l0..lc
// ---------------------------------------------------------------------------
template <class T> struct LevelDev {
  int n, ncol, ns, p_ns, t_ns;
  int gs, p_gs; // lanes per row (shift) of the level matrix and of the restriction
  int zs;       // first slice of the last colour (zero residual after a forward sweep)
  int rows_c0;  // rows of the first colour (rows [0, rows_c0))
  int rows_lc;  // first row of the last colour
  int nst;      // stored entries of the level matrix
  const int32_t *col_slice; // ncol+1 slice bounds
  const SliceInfo *si;
  const int32_t *cols;
  const T *vals, *invd, *diag;
  const SliceInfo *p_si; // restriction to level+1
  const int32_t *p_cols;
  const T *p_vals;
  const SliceInfo *t_si; // prolongation from level+1
  const int32_t *t_cols;
  const T *t_vals;
  T *x, *b, *t;
};

// The tail's operands carry their address space in the type, so that the
// compiler issues LDS (ds_*) and global (global_*) instructions rather than
// flat ones: global loads count only in the vector-memory counter, so loads
// issued before a barrier stay in flight across it (the barrier waits for
// the LDS counter), and LDS accesses take the short LDS path.

// Entries per lane loaded ahead for a colour step (most tail rows fit in
// one chunk) and the chunk of the other row products; together they keep
// the tail within 128 VGPRs.
template <class T> constexpr int TailChunk = sizeof(T) == 8 ? 6 : 12;
template <class T> constexpr int TailDotChunk = sizeof(T) == 8 ? 4 : 6;

// One wavefront's share of a colour step: the slice's first chunk of column
// and value loads, 1/diag and b of its row, issued one step ahead (none of
// them depends on x).
template <class T> struct TailRow {
  SliceInfo q;
  int32_t cc[TailChunk<T>];
  T vv[TailChunk<T>];
  T br, ir;
  bool on; // this lane has a row
};

template <class T, class SP, class VP, class MC, class MV>
__device__ __forceinline__ void tail_issue(TailRow<T> &p, const LevelDev<T> &d, SP si, VP b, int s, int lane, MC mc,
                                           MV mv) {
  p.q = load_si(si, s);
  const int gs = d.gs;
  p.on = (lane >> gs) < p.q.nrow;
  if (!p.on) return;
  const int r = p.q.row0 + (lane >> gs);
  p.br = b[r];
  p.ir = gptr(d.invd)[r];
  const auto c = mc + p.q.off + lane;
  const auto v = mv + p.q.off + lane;
#pragma unroll
  for (int k = 0; k < TailChunk<T>; k++)
    if (k < p.q.width) {
      p.cc[k] = c[k * WAVE];
      p.vv[k] = v[k * WAVE];
    }
}

// the rest of sell_dot's row product after the first chunk (same loads,
// same summation order: chunk by chunk, entry by entry)
template <class T, class XP, class MC, class MV>
__device__ __forceinline__ T tail_dot(const SliceInfo &q, int lane, MC cols, MV vals, XP x, int j0, T acc) {
  const auto c = cols + q.off + lane;
  const auto v = vals + q.off + lane;
  constexpr int CH = TailDotChunk<T>;
  for (int j = j0; j < q.width; j += CH) {
    int32_t cc[CH];
    T vv[CH], xx[CH];
#pragma unroll
    for (int k = 0; k < CH; k++)
      if (j + k < q.width) {
        cc[k] = c[(j + k) * WAVE];
        vv[k] = v[(j + k) * WAVE];
      }
#pragma unroll
    for (int k = 0; k < CH; k++)
      if (j + k < q.width) xx[k] = x[cc[k]];
#pragma unroll
    for (int k = 0; k < CH; k++)
      if (j + k < q.width) acc += vv[k] * xx[k];
  }
  return acc;
}

// x_r = (b_r - sum_{j != r} a_rj x_j) / a_rr for the row of an issued slice
template <class T, class XP, class MC, class MV>
__device__ __forceinline__ void tail_finish(const TailRow<T> &p, const LevelDev<T> &d, XP x, int lane, MC mc, MV mv) {
  if (!p.on) return;
  const int gs = d.gs;
  T acc = 0;
  {
    T xx[TailChunk<T>];
#pragma unroll
    for (int k = 0; k < TailChunk<T>; k++)
      if (k < p.q.width) xx[k] = x[p.cc[k]];
#pragma unroll
    for (int k = 0; k < TailChunk<T>; k++)
      if (k < p.q.width) acc += p.vv[k] * xx[k];
  }
  acc = tail_dot<T>(p.q, lane, mc, mv, x, TailChunk<T>, acc);
  acc = row_group_sum(acc, gs);
  if ((lane & ((1 << gs) - 1)) == 0) x[p.q.row0 + (lane >> gs)] = (p.br - acc) * p.ir;
}

// nu sweeps over the colours of a level (forward, or reverse when rev),
// wavefront w taking slices w, w + nw, ... of each colour.  The loads of a
// wavefront's first slice of the next colour are issued before the barrier
// that ends the current colour; the chunk holds a whole row's share of most
// tail levels (up to 12 entries per lane in FP32).
template <class T, class SP, class CP, class VP, class MC, class MV>
__device__ __forceinline__ void tail_sweeps(const LevelDev<T> &d, SP si, CP cs, VP b, VP x, int nu, bool rev,
                                            int w, int nw, int lane, MC mc, MV mv) {
  const int nc = d.ncol, steps = nu * nc;
  auto colour = [&](int i) {
    const int c = i % nc;
    return rev ? nc - 1 - c : c;
  };
  auto issue_step = [&](TailRow<T> &p, int i) {
    p.on = false;
    if (i < steps) {
      const int c = colour(i);
      if (cs[c] + w < cs[c + 1]) tail_issue(p, d, si, b, cs[c] + w, lane, mc, mv);
    }
  };
  TailRow<T> p;
  issue_step(p, 0);
  for (int i = 0; i < steps; i++) {
    const int c = colour(i), s1 = cs[c + 1];
    if (cs[c] + w < s1) tail_finish(p, d, x, lane, mc, mv);
    for (int s = cs[c] + w + nw; s < s1; s += nw) {
      TailRow<T> e;
      tail_issue(e, d, si, b, s, lane, mc, mv);
      tail_finish(e, d, x, lane, mc, mv);
    }
    issue_step(p, i + 1);
    __syncthreads();
  }
}

// y_r = sum_j m_rj v_j over the slices of a matrix (restriction and
// prolongation): one lane group per row, y written by the group's leader
// as f(r, sum)
template <class T, class SP, class XP, class F>
__device__ __forceinline__ void tail_rows(SP si, int ns, const int32_t *cols, const T *vals, int gs, XP v, int w,
                                          int nw, int lane, int skip0, int skip1, F f) {
  for (int s = w; s < ns; s += nw) {
    const SliceInfo q = load_si(si, s);
    const int r = q.row0 + (lane >> gs);
    if ((lane >> gs) < q.nrow && !(r >= skip0 && r < skip1)) {
      const T acc = row_group_sum(tail_dot<T>(q, lane, gptr(cols), gptr(vals), v, 0, T(0)), gs);
      if ((lane & ((1 << gs) - 1)) == 0) f(r, acc);
    }
  }
}

// LDS layout of one level while the tail works on it: [x | b | t] when the
// vectors fit (n T each), then the slice table and the colour bounds when
// they fit behind them, then the level matrix (values, int32 columns) when
// all of it fits too; otherwise those stay in global memory.  A level whose
// matrix is staged runs its colour steps on LDS alone.
struct TailStage {
  bool vec, tab, mat;
  int tab_off, mat_off; // bytes
};
template <class T> __device__ __forceinline__ TailStage tail_stage_plan(const LevelDev<T> &d, int lds_cap) {
  TailStage t;
  const int64_t vb = 3 * (int64_t)d.n * sizeof(T);
  t.vec = vb <= lds_cap;
  t.tab_off = t.vec ? (int)((vb + 15) & ~int64_t(15)) : 0;
  const int64_t te = t.tab_off + (int64_t)d.ns * sizeof(SliceInfo) + (int64_t)(d.ncol + 1) * 4;
  t.tab = te <= lds_cap;
  // the level matrix itself (values, then int32 columns) when everything fits
  t.mat_off = (int)((te + 15) & ~int64_t(15));
  t.mat = t.vec && t.tab && t.mat_off + (int64_t)d.nst * (sizeof(T) + 4) <= lds_cap;
  return t;
}

// Single-workgroup V-cycle over the small levels l0..lc.  phase bits:
// 1 down leg (pre-smoothing, residual, restriction for l0..lc-1),
// 2 coarse solve on level lc (dense FP64 inverse, diagonal or GS),
// 4 up leg (prolongation, post-smoothing for lc-1..l0).
// lds_cap bytes of dynamic LDS hold the working level's vectors and tables
// (tail_stage_plan); x of a level is written back to global memory at the
// end of each leg.
template <class T, class SP, class CP, class VP, class MC, class MV>
__device__ __forceinline__ void tail_down(const LevelDev<T> &d, const LevelDev<T> &dn, SP si, CP cs, VP x, VP b,
                                          VP t, int nu, int w, int nw, int lane, MC mc, MV mv) {
  tail_sweeps<T>(d, si, cs, b, x, nu, false, w, nw, lane, mc, mv);
  // t = b - A x (zero on the last colour after a sweep, as k_residual)
  const int s_zero = nu > 0 ? d.zs : d.ns;
  for (int s = w; s < d.ns; s += nw) {
    const SliceInfo q = load_si(si, s);
    const int gs = d.gs;
    if ((lane >> gs) < q.nrow) {
      const int r = q.row0 + (lane >> gs);
      if (s >= s_zero) {
        if ((lane & ((1 << gs) - 1)) == 0) t[r] = 0;
        continue;
      }
      const T br = b[r], dr = gptr(d.diag)[r], xr = x[r];
      const T acc = row_group_sum(tail_dot<T>(q, lane, mc, mv, x, 0, T(0)), gs);
      if ((lane & ((1 << gs) - 1)) == 0) t[r] = br - dr * xr - acc;
    }
  }
  __syncthreads();
  // b_{l+1} = P t, x_{l+1} = 0
  const GPtr<T> nb = gptr(dn.b), nx = gptr(dn.x);
  tail_rows<T>(gptr(d.p_si), d.p_ns, d.p_cols, d.p_vals, d.p_gs, t, w, nw, lane, 0, 0, [&](int r, T v) {
    nb[r] = v;
    nx[r] = 0;
  });
}

template <class T, class SP, class CP, class VP, class MC, class MV>
__device__ __forceinline__ void tail_up(const LevelDev<T> &d, const LevelDev<T> &dn, SP si, CP cs, VP x, VP b,
                                        int nu, bool rev, int w, int nw, int lane, MC mc, MV mv) {
  // x += P^T x_{l+1}, except on the colour the post-smoother sweeps first
  // (overwritten unread, as in k_prolong)
  const GPtr<const T> nx = gptr((const T *)dn.x);
  const int skip0 = nu == 0 ? 0 : (rev ? d.rows_lc : 0), skip1 = nu == 0 ? 0 : (rev ? d.n : d.rows_c0);
  tail_rows<T>(gptr(d.t_si), d.t_ns, d.t_cols, d.t_vals, 0, nx, w, nw, lane, skip0, skip1,
               [&](int r, T v) { x[r] = x[r] + v; });
  __syncthreads();
  tail_sweeps<T>(d, si, cs, b, x, nu, rev, w, nw, lane, mc, mv);
}

// runs f(si, cs, x, b, t) with the level's tables and vectors in LDS or in
// global memory as planned (staging copies made and synchronised first)
template <class T, class F>
__device__ __forceinline__ void tail_with_level(const LevelDev<T> &d, char *smem, int lds_cap, F f) {
  const TailStage pl = tail_stage_plan(d, lds_cap);
  const int tid = threadIdx.x;
  LPtr<T> sx = (LPtr<T>)(T *)smem;
  LPtr<T> sb = sx + d.n, st = sb + d.n;
  LPtr<SliceInfo> ssi = (LPtr<SliceInfo>)(SliceInfo *)(smem + pl.tab_off);
  LPtr<int32_t> scs = (LPtr<int32_t>)(int32_t *)(smem + pl.tab_off + (int64_t)d.ns * sizeof(SliceInfo));
  LPtr<T> smv = (LPtr<T>)(T *)(smem + pl.mat_off);
  LPtr<int32_t> smc = (LPtr<int32_t>)(int32_t *)(smem + pl.mat_off + (int64_t)d.nst * sizeof(T));
  if (pl.mat)
    for (int i = tid; i < d.nst; i += blockDim.x) {
      smv[i] = gptr(d.vals)[i];
      smc[i] = gptr(d.cols)[i];
    }
  if (pl.vec)
    for (int i = tid; i < d.n; i += blockDim.x) {
      sx[i] = gptr(d.x)[i];
      sb[i] = gptr(d.b)[i];
    }
  if (pl.tab) {
    for (int i = tid; i < d.ns; i += blockDim.x) {
      const SliceInfo q = d.si[i];
      ssi[i].row0 = q.row0;
      ssi[i].nrow = q.nrow;
      ssi[i].width = q.width;
      ssi[i].off = q.off;
    }
    for (int i = tid; i <= d.ncol; i += blockDim.x) scs[i] = gptr(d.col_slice)[i];
  }
  __syncthreads();
  const GPtr<const int32_t> gc = gptr(d.cols);
  const GPtr<const T> gv = gptr(d.vals);
  if (pl.vec) {
    if (pl.mat)
      f(ssi, scs, sx, sb, st, (LPtr<const int32_t>)smc, (LPtr<const T>)smv);
    else if (pl.tab)
      f(ssi, scs, sx, sb, st, gc, gv);
    else
      f(gptr(d.si), gptr(d.col_slice), sx, sb, st, gc, gv);
    for (int i = tid; i < d.n; i += blockDim.x) gptr(d.x)[i] = sx[i];
  } else if (pl.tab) {
    f(ssi, scs, gptr(d.x), gptr(d.b), gptr(d.t), gc, gv);
  } else {
    f(gptr(d.si), gptr(d.col_slice), gptr(d.x), gptr(d.b), gptr(d.t), gc, gv);
  }
  __syncthreads();
}

template <class T>
__device__ __forceinline__ void tail_body(const LevelDev<T> *__restrict__ L, int l0, int lc,
                                          const double *__restrict__ minv, int coarse_mode, int nu_pre,
                                          int nu_post, int post_reverse, int phase, int lds_cap) {
  extern __shared__ __align__(16) char tail_smem[];
  const int nw = blockDim.x / WAVE;
  const int w = wave_id();
  const int lane = threadIdx.x % WAVE;
  const int tid = threadIdx.x;
  if (phase & 1) {
    for (int l = l0; l < lc; l++) {
      const LevelDev<T> &d = L[l];
      const LevelDev<T> &dn = L[l + 1];
      tail_with_level<T>(d, tail_smem, lds_cap, [&](auto si, auto cs, auto x, auto b, auto t, auto mc, auto mv) {
        tail_down<T>(d, dn, si, cs, x, b, t, nu_pre, w, nw, lane, mc, mv);
      });
    }
  }
  if (phase & 2) {
    const LevelDev<T> &d = L[lc];
    if (coarse_mode == 0) {
      const GPtr<const T> b = gptr((const T *)d.b);
      const GPtr<T> x = gptr(d.x);
      for (int i = tid; i < d.n; i += blockDim.x) {
        const GPtr<const double> m = gptr(minv) + (int64_t)i * d.n;
        double s = 0.0;
        for (int j0 = 0; j0 < d.n; j0 += 8) { // 8 loads in flight, summed in order
          double mm[8];
          T bb[8];
#pragma unroll
          for (int k = 0; k < 8; k++)
            if (j0 + k < d.n) {
              mm[k] = m[j0 + k];
              bb[k] = b[j0 + k];
            }
#pragma unroll
          for (int k = 0; k < 8; k++)
            if (j0 + k < d.n) s += mm[k] * (double)bb[k];
        }
        x[i] = (T)s;
      }
    } else if (coarse_mode == 1) {
      for (int i = tid; i < d.n; i += blockDim.x) gptr(d.x)[i] = gptr(d.b)[i] * gptr(d.invd)[i];
    } else {
      tail_sweeps<T>(d, gptr(d.si), gptr(d.col_slice), gptr(d.b), gptr(d.x), nu_pre, false, w, nw, lane,
                     gptr(d.cols), gptr(d.vals));
    }
    __syncthreads();
  }
  if (phase & 4) {
    for (int l = lc - 1; l >= l0; l--) {
      const LevelDev<T> &d = L[l];
      const LevelDev<T> &dn = L[l + 1];
      tail_with_level<T>(d, tail_smem, lds_cap, [&](auto si, auto cs, auto x, auto b, auto, auto mc, auto mv) {
        tail_up<T>(d, dn, si, cs, x, b, nu_post, post_reverse != 0, w, nw, lane, mc, mv);
      });
    }
  }
}

template <class T>
__global__ __launch_bounds__(TAIL_THREADS) void k_tail(const LevelDev<T> *__restrict__ L, int l0, int lc,
                                                       const double *__restrict__ minv, int coarse_mode,
                                                       int nu_pre, int nu_post, int post_reverse, int phase,
                                                       int lds_cap) {
  tail_body<T>(L, l0, lc, minv, coarse_mode, nu_pre, nu_post, post_reverse, phase, lds_cap);
}

// dense coarse solve x = Minv b with the FP64 inverse (DenseMatrix::vmult of
// the coarsest level, src/AMGSolver.cpp:591-593): one wavefront per row,
// coalesced row reads, FP64 accumulation; for coarse levels too large for
// the single-workgroup tail
template <class T>
__global__ __launch_bounds__(BLK) void k_coarse_gemv(int n, const double *__restrict__ minv,
                                                     const T *__restrict__ b, T *__restrict__ x) {
  __shared__ double sh[BLK];
  const int row = blockIdx.x * WPB + wave_id();
  const int lane = threadIdx.x % WAVE;
  double s = 0.0;
  if (row < n) {
    // loads of 8 strides issued together, then summed in the same order
    const double *m = minv + (int64_t)row * n;
    for (int j0 = lane; j0 < n; j0 += 8 * WAVE) {
      double mm[8];
      T bb[8];
#pragma unroll
      for (int k = 0; k < 8; k++)
        if (j0 + k * WAVE < n) {
          mm[k] = m[j0 + k * WAVE];
          bb[k] = b[j0 + k * WAVE];
        }
#pragma unroll
      for (int k = 0; k < 8; k++)
        if (j0 + k * WAVE < n) s += mm[k] * (double)bb[k];
    }
  }
  // wavefront reduction through LDS (fixed order)
  sh[threadIdx.x] = s;
  __syncthreads();
  const int base = threadIdx.x - lane;
  for (int o = WAVE / 2; o > 0; o >>= 1) {
    if (lane < o) sh[base + lane] += sh[base + lane + o];
    __syncthreads();
  }
  if (row < n && lane == 0) x[row] = (T)sh[base];
}

// ---------------------------------------------------------------------------
// vector kernels (level-0 numbering after the colour permutation)
// ---------------------------------------------------------------------------
// plain streaming copy (16-byte vectors, grid-stride): the calibration
// kernel of the byte counters and of the achievable HBM bandwidth
typedef double dvec2 __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(256) void k_stream_copy(int64_t n4, const dvec2 *__restrict__ src,
                                                     dvec2 *__restrict__ dst) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride)
    __builtin_nontemporal_store(__builtin_nontemporal_load(&src[i]), &dst[i]);
}

template <class T>
__global__ __launch_bounds__(256) void k_fill(int64_t n, T *__restrict__ x, T v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

template <class S, class D>
__global__ __launch_bounds__(256) void k_permute_in(int64_t n, const int32_t *__restrict__ perm,
                                                    const S *__restrict__ src, D *__restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (D)src[perm[i]];
}

template <class S, class D>
__global__ __launch_bounds__(256) void k_permute_out(int64_t n, const int32_t *__restrict__ perm,
                                                     const S *__restrict__ src, D *__restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[perm[i]] = (D)src[i];
}

// y = A x (FP64, level 0) and the partial dots (x, y) per block; grid-stride
// over slices with a fixed grid so the partial count is RED_BLOCKS
__global__ __launch_bounds__(BLK) void k_spmv_dot(const SliceInfo *__restrict__ si, int ns,
                                                  const int32_t *__restrict__ cols,
                                                  const uint16_t *__restrict__ c16,
                                                  const int32_t *__restrict__ cbase,
                                                  const double *__restrict__ vals,
                                                  const double *__restrict__ diag,
                                                  const int32_t *__restrict__ rowmap,
                                                  const double *__restrict__ x, double *__restrict__ y,
                                                  double *__restrict__ part, int xcd) {
  __shared__ double sh[BLK];
  const int lane = threadIdx.x % WAVE;
  double dot = 0.0;
  int s0, s_end, step;
  slice_range(ns, xcd, s0, s_end, step);
  for (int s = s0; s < s_end; s += step) {
    const SliceInfo q = si[s];
    if (lane < q.nrow) {
      const int pos = q.row0 + lane;
      const int r = rowmap ? rowmap[pos] : pos;
      const double xr = x[r];
      const double acc = c16 ? sell_dot<double, true, SELL_CHUNK_SPMV, uint16_t>(q, lane, c16, vals, x, cbase[s])
                             : sell_dot<double, true, SELL_CHUNK_SPMV>(q, lane, cols, vals, x);
      const double yr = diag[pos] * xr + acc;
      y[r] = yr;
      dot += xr * yr;
    }
  }
  dot = block_sum(dot, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = dot;
}

// ---------------------------------------------------------------------------
// value-only refresh of a lazy hierarchy (AMGSolver::lazyReinit,
// src/AMGSolver.cpp:505-545): the Galerkin products with the kept P on the
// fixed coarse patterns, then the layouts' values from the CSR values
// ---------------------------------------------------------------------------
#pragma clang fp contract(off)
// one coarse row per thread: C(i, :) = sum_j P(i, a) A(a, b) P^T(b, :), the
// terms accumulated in the order of the host product (P row, A row, P^T row)
// so that every coarse value is bit-identical to the host's; the coarse row
// holds the diagonal first and the other columns ascending
__global__ __launch_bounds__(256) void k_galerkin_rows(
    int64_t nc, const int64_t *__restrict__ Prp, const int32_t *__restrict__ Pcol,
    const double *__restrict__ Pval, const int64_t *__restrict__ Arp, const int32_t *__restrict__ Acol,
    const double *__restrict__ Aval, const int64_t *__restrict__ Ptrp, const int32_t *__restrict__ Ptcol,
    const double *__restrict__ Ptval, const int64_t *__restrict__ Crp, const int32_t *__restrict__ Ccol,
    double *__restrict__ Cval) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc) return;
  const int64_t c0 = Crp[i], c1 = Crp[i + 1];
  for (int64_t k = c0; k < c1; k++) Cval[k] = 0.0;
  for (int64_t j = Prp[i]; j < Prp[i + 1]; j++) {
    const int64_t a = Pcol[j];
    const double pa = Pval[j];
    for (int64_t k = Arp[a]; k < Arp[a + 1]; k++) {
      const int64_t b = Acol[k];
      const double pab = pa * Aval[k];
      for (int64_t l = Ptrp[b]; l < Ptrp[b + 1]; l++) {
        const int32_t cc = Ptcol[l];
        int64_t pos = -1;
        if (cc == i) {
          pos = c0;
        } else {
          int64_t lo = c0 + 1, hi = c1 - 1;
          while (lo <= hi) {
            const int64_t mid = (lo + hi) >> 1;
            const int32_t v = Ccol[mid];
            if (v == cc) {
              pos = mid;
              break;
            }
            if (v < cc) lo = mid + 1;
            else hi = mid - 1;
          }
        }
        if (pos >= 0) Cval[pos] += pab * Ptval[l];
      }
    }
  }
}

// layout values from CSR values through the source map (-1: split diagonal)
__global__ __launch_bounds__(256) void k_refresh_vals(int64_t nnz, const double *__restrict__ val,
                                                      const int32_t *__restrict__ smap, double *__restrict__ v64,
                                                      float *__restrict__ v32) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const int32_t p = smap[k];
  if (p < 0) return;
  v64[p] = val[k];
  if (v32) v32[p] = (float)val[k];
}

__global__ __launch_bounds__(256) void k_refresh_diag(int64_t n, const double *__restrict__ val,
                                                      const int32_t *__restrict__ dsrc, double *__restrict__ diag64,
                                                      double *__restrict__ invd64, float *__restrict__ diag32,
                                                      float *__restrict__ invd32) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const double d = val[dsrc[r]];
  diag64[r] = d;
  if (invd64) invd64[r] = 1.0 / d;
  if (diag32) diag32[r] = (float)d;
  if (invd32) invd32[r] = (float)(1.0 / d);
}
#pragma clang fp contract(on)

// error-free transformations (no contraction inside: fp contract is off in
// this region; the product error uses an explicit fma)
#pragma clang fp contract(off)
__device__ inline void two_sum(double a, double b, double &s, double &e) {
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// r = b - A (x + xl) evaluated with compensated (error-free) products and
// sums, and the partials of (r, r), (b, b), sum |r|; xl may be null
__global__ __launch_bounds__(BLK) void k_residual_norms_dd(const SliceInfo *__restrict__ si, int ns,
                                                           const int32_t *__restrict__ cols,
                                                           const uint16_t *__restrict__ c16,
                                                           const int32_t *__restrict__ cbase,
                                                           const double *__restrict__ vals,
                                                           const double *__restrict__ diag,
                                                           const int32_t *__restrict__ rowmap,
                                                           const double *__restrict__ b,
                                                           const double *__restrict__ x,
                                                           const double *__restrict__ xl,
                                                           double *__restrict__ r, double *__restrict__ part,
                                                           int xcd) {
  __shared__ double sh[BLK];
  const int lane = threadIdx.x % WAVE;
  double rr = 0.0, bb = 0.0, l1 = 0.0;
  int s0, s_end, step;
  slice_range(ns, xcd, s0, s_end, step);
  for (int s = s0; s < s_end; s += step) {
    const SliceInfo q = si[s];
    if (lane < q.nrow) {
      const int pos = q.row0 + lane;
      const int row = rowmap ? rowmap[pos] : pos;
      double hi = b[row], lo = 0.0;
      {
        const double a = diag[pos], xv = x[row];
        const double pr = a * xv, pe = fma(a, xv, -pr);
        double t, e;
        two_sum(hi, -pr, t, e);
        hi = t;
        lo += e - pe;
        if (xl) lo -= a * xl[row];
      }
      const int32_t *c = c16 ? nullptr : cols + q.off + lane;
      const uint16_t *c2 = c16 ? c16 + q.off + lane : nullptr;
      const int32_t cb = c16 ? cbase[s] : 0;
      const double *v = vals + q.off + lane;
      // chunks of CH entries: the column and value loads, then the gathers,
      // then the compensated products in stored order (as one at a time)
      constexpr int CH = SELL_CHUNK_SPMV;
      for (int j0 = 0; j0 < q.width; j0 += CH) {
        int32_t cc[CH];
        double vv[CH], xx[CH], xe[CH];
#pragma unroll
        for (int k = 0; k < CH; k++)
          if (j0 + k < q.width) {
            cc[k] = c16 ? cb + (int32_t)c2[(j0 + k) * WAVE] : c[(j0 + k) * WAVE];
            vv[k] = v[(j0 + k) * WAVE];
          }
#pragma unroll
        for (int k = 0; k < CH; k++)
          if (j0 + k < q.width) {
            xx[k] = x[cc[k]];
            xe[k] = xl ? xl[cc[k]] : 0.0;
          }
#pragma unroll
        for (int k = 0; k < CH; k++)
          if (j0 + k < q.width) {
            const double a = vv[k], xv = xx[k];
            const double pr = a * xv, pe = fma(a, xv, -pr);
            double t, e;
            two_sum(hi, -pr, t, e);
            hi = t;
            lo += e - pe;
            if (xl) lo -= a * xe[k];
          }
      }
      const double rv = hi + lo;
      if (r) r[row] = rv;
      rr += rv * rv;
      bb += b[row] * b[row];
      l1 += fabs(rv);
    }
  }
  rr = block_sum(rr, sh);
  bb = block_sum(bb, sh);
  l1 = block_sum(l1, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = rr;
    part[gridDim.x + blockIdx.x] = bb;
    part[2 * gridDim.x + blockIdx.x] = l1;
  }
}

// x = round(x + xl), xl = 0
__global__ __launch_bounds__(256) void k_dd_round(int64_t n, double *__restrict__ x, double *__restrict__ xl) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    x[i] = x[i] + xl[i];
    xl[i] = 0.0;
  }
}
#pragma clang fp contract(on)

// device scalars of the CG loop; k_finalize writes consecutive fields
// (rz_new, rz_cross) and (rr, bb, l1) in one call
struct CgScalars {
  double rz, pq, rz_new, rz_cross, rr, bb, l1, beta;
};

// k_finalize of one sum into *out (a field of *sc), then the scalars copied
// to pinned host memory by the thread that wrote the sum, so the host reads
// them once the launch is complete without a separate copy (vector stores
// to host-coherent memory; the launch's completion makes them visible)
__global__ __launch_bounds__(FIN_THREADS) void k_finalize_host(const double *__restrict__ part, int np,
                                                               double *out, const CgScalars *sc,
                                                               CgScalars *__restrict__ h_sc) {
  __shared__ double sh[FIN_THREADS];
  double s = strided_sum(part, np, threadIdx.x, blockDim.x);
  s = block_sum(s, sh);
  if (threadIdx.x == 0) {
    *out = s;
    const CgScalars v = *sc;
    *h_sc = v;
  }
}

#pragma clang fp contract(off)
// alpha = rz / pq; x += alpha p kept as the unevaluated pair x_hi + x_lo
// (the sum and the product errors go to x_lo); r -= alpha q;
// (T) r -> b0, x0 = 0; partial (r, r)
template <class T>
__global__ __launch_bounds__(256) void k_cg_update(int64_t n, const CgScalars *__restrict__ sc,
                                                   double *__restrict__ x, double *__restrict__ xl,
                                                   double *__restrict__ r,
                                                   const double *__restrict__ p, const double *__restrict__ q,
                                                   T *__restrict__ b0, T *__restrict__ x0,
                                                   const int32_t *__restrict__ iperm,
                                                   double *__restrict__ part) {
  __shared__ double sh[256];
  const double alpha = sc->rz / sc->pq;
  double rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double ap = alpha * p[i];
    const double ape = fma(alpha, p[i], -ap);
    double s, e;
    two_sum(x[i], ap, s, e);
    x[i] = s;
    xl[i] += e + ape;
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    b0[iperm ? iperm[i] : i] = (T)ri;
    if (x0) x0[i] = 0; // null: the V-cycle's first colour kernel sets x0
    rr += ri * ri;
  }
  rr = block_sum(rr, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = rr;
}

#pragma clang fp contract(on)

// (T) r -> b0 (colour order through iperm when given), x0 = 0 (start of a
// preconditioner application)
template <class T>
__global__ __launch_bounds__(256) void k_prep_vcycle(int64_t n, const double *__restrict__ r,
                                                     T *__restrict__ b0, T *__restrict__ x0,
                                                     const int32_t *__restrict__ iperm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    b0[iperm ? iperm[i] : i] = (T)r[i];
    x0[i] = 0;
  }
}

// z_new = (double) x0 (gathered from colour order through iperm when given);
// partials (r, z_new) and (r, z_old); z = z_new
template <class T>
__global__ __launch_bounds__(256) void k_take_z(int64_t n, const T *__restrict__ x0,
                                                const double *__restrict__ r, double *__restrict__ z,
                                                const int32_t *__restrict__ iperm,
                                                double *__restrict__ part) {
  __shared__ double sh[256];
  double a = 0.0, b = 0.0;
  // four grid strides at a time: the index loads, then the gathers, then the
  // sums in stride order (the same sums as one stride at a time)
  constexpr int U = 4;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i0 < n; i0 += U * stride) {
    int32_t ip[U];
    double rv[U], zo[U];
    T xv[U];
#pragma unroll
    for (int k = 0; k < U; k++) {
      const int64_t i = i0 + k * stride;
      if (i < n) {
        ip[k] = iperm ? iperm[i] : (int32_t)i;
        rv[k] = r[i];
        zo[k] = z[i];
      }
    }
#pragma unroll
    for (int k = 0; k < U; k++)
      if (i0 + k * stride < n) xv[k] = x0[ip[k]];
#pragma unroll
    for (int k = 0; k < U; k++) {
      const int64_t i = i0 + k * stride;
      if (i < n) {
        const double zn = (double)xv[k];
        a += rv[k] * zn;
        b += rv[k] * zo[k];
        z[i] = zn;
      }
    }
  }
  a = block_sum(a, sh);
  b = block_sum(b, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[gridDim.x + blockIdx.x] = b;
  }
}

// k_finalize of the (r, z_new) and (r, z_old) partials into rz_new and
// rz_cross, then (the same thread, in order) the flexible (Polak-Ribiere)
// beta = (rz_new - rz_cross) / rz, robust to the FP32 preconditioner
// (first != 0: beta = 0, p = z), and rz = rz_new
__global__ __launch_bounds__(FIN_THREADS) void k_finalize_beta(const double *__restrict__ part, int np,
                                                               CgScalars *__restrict__ sc, int first) {
  __shared__ double sh[FIN_THREADS];
  double v[2];
  for (int k = 0; k < 2; k++) v[k] = block_sum(strided_sum(part + (int64_t)k * np, np, threadIdx.x, blockDim.x), sh);
  if (threadIdx.x == 0) {
    sc->rz_new = v[0];
    sc->rz_cross = v[1];
    sc->beta = first ? 0.0 : (v[0] - v[1]) / sc->rz;
    sc->rz = v[0];
  }
}

__global__ __launch_bounds__(256) void k_p_update(int64_t n, const CgScalars *__restrict__ sc,
                                                  const double *__restrict__ z, double *__restrict__ p) {
  const double beta = sc->beta;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = z[i] + beta * p[i];
}

} // namespace afe
