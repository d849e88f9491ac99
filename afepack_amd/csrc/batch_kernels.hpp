// batch_kernels.hpp -- one launch for the same step of many independent
// solves (continuous batching of the AMG-preconditioned CG).
//
// A batched launch covers a list of segments; a segment is one member's share
// of the launch: a contiguous run of blocks working on a contiguous range of
// that member's slices (or rows).  Blocks never mix members, and a member's
// slices are visited by its own blocks in the same way whatever else is in
// the batch, so every member's result -- reductions included, which are
// summed per member in block order -- is bitwise the same alone and in any
// mix.
#pragma once
#include "kernels.hpp"

namespace afe {

struct Seg {
  int32_t block0; // first block of the segment in the grid
  int32_t member; // index into the launch's member tables
  int32_t s0, s1; // slice (or row) range of the member
};

// the segment of this block (segments sorted by block0, block0[0] = 0)
__device__ inline Seg find_seg(const Seg *__restrict__ segs, int nseg) {
  int lo = 0, hi = nseg - 1;
  const int b = blockIdx.x;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (segs[mid].block0 <= b) lo = mid;
    else hi = mid - 1;
  }
  return segs[lo];
}

// per-member data of the CG in the natural numbering and the level-0 vectors
// of the V-cycle (colour order)
template <class T> struct MemberDev {
  int64_t n;
  const SliceInfo *nat_si;
  const int32_t *nat_cols, *nat_rowmap, *iperm0;
  const uint16_t *nat_c16; // 16-bit columns (offsets from nat_cbase per slice) when not null
  const int32_t *nat_cbase;
  const double *nat_v, *nat_diag;
  double *x, *xl, *r, *p, *q, *z;
  T *b0, *x0;
  CgScalars *sc;
};

// ---- V-cycle level kernels: one wavefront per slice, 2^gs lanes per row ---
// A block covers tile * WPB consecutive slices of its segment: wavefront w
// takes slices w, w + WPB, ... (tile of them).  The rows of one launch are
// independent, so the results do not depend on the tile.
template <class T>
__global__ __launch_bounds__(BLK) void kb_gs(const Seg *__restrict__ segs, int nseg,
                                             const LevelDev<T> *__restrict__ lv, int tile) {
  const Seg g = find_seg(segs, nseg);
  const LevelDev<T> &d = lv[g.member];
  const int lane = threadIdx.x % WAVE;
  const int gs = d.gs;
  for (int t = 0; t < tile; t++) {
    const int s = g.s0 + ((blockIdx.x - g.block0) * tile + t) * WPB + wave_id();
    if (s >= g.s1) return;
    const SliceInfo q = load_si(gptr(d.si), s);
    if ((lane >> gs) >= q.nrow) continue;
    const int r = q.row0 + (lane >> gs);
    const T br = gptr(d.b)[r], ir = gptr(d.invd)[r];
    const T acc = row_group_sum(sell_dot<T>(q, lane, gptr(d.cols), gptr(d.vals), gptr((const T *)d.x)), gs);
    if ((lane & ((1 << gs) - 1)) == 0) gptr(d.x)[r] = (br - acc) * ir;
  }
}

template <class T>
__global__ __launch_bounds__(BLK) void kb_residual(const Seg *__restrict__ segs, int nseg,
                                                   const LevelDev<T> *__restrict__ lv, int tile) {
  const Seg g = find_seg(segs, nseg);
  const LevelDev<T> &d = lv[g.member];
  const int lane = threadIdx.x % WAVE;
  const int gs = d.gs;
  for (int t = 0; t < tile; t++) {
    const int s = g.s0 + ((blockIdx.x - g.block0) * tile + t) * WPB + wave_id();
    if (s >= g.s1) return;
    const SliceInfo q = load_si(gptr(d.si), s);
    if ((lane >> gs) >= q.nrow) continue;
    const int r = q.row0 + (lane >> gs);
    if (s >= d.zs) { // the last colour after the forward sweep (the batch smooths at least once)
      if ((lane & ((1 << gs) - 1)) == 0) gptr(d.t)[r] = 0;
      continue;
    }
    const T br = gptr(d.b)[r], dr = gptr(d.diag)[r], xr = gptr(d.x)[r];
    const T acc = row_group_sum(sell_dot<T>(q, lane, gptr(d.cols), gptr(d.vals), gptr((const T *)d.x)), gs);
    if ((lane & ((1 << gs) - 1)) == 0) gptr(d.t)[r] = br - dr * xr - acc;
  }
}

// b_{l+1} = P t_l, x_{l+1} = 0 (lv: level l, lvn: level l+1 of each member)
template <class T>
__global__ __launch_bounds__(BLK) void kb_restrict(const Seg *__restrict__ segs, int nseg,
                                                   const LevelDev<T> *__restrict__ lv,
                                                   const LevelDev<T> *__restrict__ lvn, int tile) {
  const Seg g = find_seg(segs, nseg);
  const LevelDev<T> &d = lv[g.member];
  const LevelDev<T> &dn = lvn[g.member];
  const int lane = threadIdx.x % WAVE;
  const int gs = d.p_gs;
  for (int t = 0; t < tile; t++) {
    const int s = g.s0 + ((blockIdx.x - g.block0) * tile + t) * WPB + wave_id();
    if (s >= g.s1) return;
    const SliceInfo q = load_si(gptr(d.p_si), s);
    if ((lane >> gs) >= q.nrow) continue;
    const int r = q.row0 + (lane >> gs);
    const T acc =
        row_group_sum(sell_dot<T>(q, lane, gptr(d.p_cols), gptr(d.p_vals), gptr((const T *)d.t)), gs);
    if ((lane & ((1 << gs) - 1)) == 0) {
      // the next level's first colour from zero (as k_restrict); a coarse
      // tail recomputes the same values when it sweeps that colour
      gptr(dn.b)[r] = acc;
      gptr(dn.x)[r] = r < dn.rows_c0 ? acc * gptr(dn.invd)[r] : (T)0;
    }
  }
}

// x_l += P^T x_{l+1}
template <class T>
__global__ __launch_bounds__(BLK) void kb_prolong(const Seg *__restrict__ segs, int nseg,
                                                  const LevelDev<T> *__restrict__ lv,
                                                  const LevelDev<T> *__restrict__ lvn, int tile) {
  const Seg g = find_seg(segs, nseg);
  const LevelDev<T> &d = lv[g.member];
  const LevelDev<T> &dn = lvn[g.member];
  const int lane = threadIdx.x % WAVE;
  for (int t = 0; t < tile; t++) {
    const int s = g.s0 + ((blockIdx.x - g.block0) * tile + t) * WPB + wave_id();
    if (s >= g.s1) return;
    const SliceInfo q = load_si(gptr(d.t_si), s);
    if (lane >= q.nrow) continue;
    const int r = q.row0 + lane;
    if (r >= d.rows_lc) continue; // the last colour: the reverse post-sweep overwrites it unread
    const T xr = gptr(d.x)[r];
    gptr(d.x)[r] = xr + sell_dot<T>(q, lane, gptr(d.t_cols), gptr(d.t_vals), gptr((const T *)dn.x));
  }
}

// the coarse tail of each member: one workgroup per member
// a member's level descriptors are at tl + loff in one table of the batch:
// reached from a kernel argument, their fields load as scalars (a table
// pointer read from memory would make every field a flat vector load)
template <class T> struct TailJob {
  const double *minv;
  int loff, l0, lc, coarse_mode;
};

template <class T>
__global__ __launch_bounds__(TAIL_THREADS) void kb_tail(const TailJob<T> *__restrict__ jobs,
                                                        const LevelDev<T> *__restrict__ tl, int nu,
                                                        int lds_cap) {
  const TailJob<T> j = jobs[blockIdx.x];
  tail_body<T>(tl + j.loff, j.l0, j.lc, j.minv, j.coarse_mode, nu, nu, 1, 7, lds_cap);
}

// ---- CG kernels (natural numbering) ---------------------------------------
// q = A p and the block partial of (p, q); segments over natural slices
template <class T>
__global__ __launch_bounds__(BLK) void kb_spmv_dot(const Seg *__restrict__ segs, int nseg,
                                                   const MemberDev<T> *__restrict__ md,
                                                   double *__restrict__ part) {
  __shared__ double sh[BLK];
  const Seg g = find_seg(segs, nseg);
  const int s = g.s0 + (blockIdx.x - g.block0) * WPB + wave_id();
  const int lane = threadIdx.x % WAVE;
  double dot = 0.0;
  if (s < g.s1) {
    const MemberDev<T> &m = md[g.member];
    const SliceInfo q = load_si(gptr(m.nat_si), s);
    if (lane < q.nrow) {
      const int pos = q.row0 + lane;
      const int r = gptr(m.nat_rowmap)[pos];
      const GPtr<const double> p = gptr((const double *)m.p);
      const double xr = p[r];
      const double acc =
          m.nat_c16 ? sell_dot<double, false, SELL_CHUNK_SPMV>(q, lane, gptr(m.nat_c16), gptr(m.nat_v), p,
                                                               gptr(m.nat_cbase)[s])
                    : sell_dot<double, false, SELL_CHUNK_SPMV>(q, lane, gptr(m.nat_cols), gptr(m.nat_v), p);
      const double yr = gptr(m.nat_diag)[pos] * xr + acc;
      gptr(m.q)[r] = yr;
      dot = xr * yr;
    }
  }
  dot = block_sum(dot, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = dot;
}

// sum of each member's block partials, in block order, into a scalar slot:
// one block per member; first[m], count[m] = the member's block range
// with beta set, the (rz_new, rz_cross) sums are followed (same thread, in
// order) by the member's flexible beta and rz = rz_new (rz = 0 on a member's
// first step gives beta = 0)
__global__ __launch_bounds__(256) void kb_finalize(const int32_t *__restrict__ first,
                                                   const int32_t *__restrict__ count, int nsum, int stride,
                                                   const double *__restrict__ part,
                                                   CgScalars *const *__restrict__ sc, int field, int beta = 0) {
  __shared__ double sh[256];
  const int m = blockIdx.x;
  for (int k = 0; k < nsum; k++) {
    double s = strided_sum(part + (int64_t)k * stride + first[m], count[m], threadIdx.x, blockDim.x);
    s = block_sum(s, sh);
    if (threadIdx.x == 0) (&sc[m]->rz)[field + k] = s;
  }
  if (beta && threadIdx.x == 0) {
    CgScalars *c = sc[m];
    c->beta = c->rz != 0.0 ? (c->rz_new - c->rz_cross) / c->rz : 0.0;
    c->rz = c->rz_new;
  }
}

#pragma clang fp contract(off)
// x += alpha p (double-double pair), r -= alpha q, (T) r -> b0 (colour
// order), x0 = 0, block partial of (r, r); segments over rows, 256 per block
template <class T>
__global__ __launch_bounds__(256) void kb_cg_update(const Seg *__restrict__ segs, int nseg,
                                                    const MemberDev<T> *__restrict__ md,
                                                    double *__restrict__ part) {
  __shared__ double sh[256];
  const Seg g = find_seg(segs, nseg);
  const int64_t i = (int64_t)g.s0 + (int64_t)(blockIdx.x - g.block0) * 256 + threadIdx.x;
  double rr = 0.0;
  if (i < g.s1) {
    const MemberDev<T> &m = md[g.member];
    // p = 0 (a member's first step after its start or a restart): pq = 0
    // and alpha = 0, so x and r stay
    const double pq = m.sc->pq;
    const double alpha = pq != 0.0 ? m.sc->rz / pq : 0.0;
    const double pi = gptr(m.p)[i];
    const double ap = alpha * pi;
    const double ape = fma(alpha, pi, -ap);
    double s, e;
    two_sum(gptr(m.x)[i], ap, s, e);
    gptr(m.x)[i] = s;
    gptr(m.xl)[i] += e + ape;
    const double ri = gptr(m.r)[i] - alpha * gptr(m.q)[i];
    gptr(m.r)[i] = ri;
    gptr(m.b0)[gptr(m.iperm0)[i]] = (T)ri;
    gptr(m.x0)[i] = 0;
    rr = ri * ri;
  }
  rr = block_sum(rr, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = rr;
}
#pragma clang fp contract(on)

// z_new = (double) x0 gathered from colour order; partials (r, z_new) and
// (r, z_old) at part[b] and part[stride + b]; z = z_new
template <class T>
__global__ __launch_bounds__(256) void kb_take_z(const Seg *__restrict__ segs, int nseg,
                                                 const MemberDev<T> *__restrict__ md, int stride,
                                                 double *__restrict__ part) {
  __shared__ double sh[256];
  const Seg g = find_seg(segs, nseg);
  const int64_t i = (int64_t)g.s0 + (int64_t)(blockIdx.x - g.block0) * 256 + threadIdx.x;
  double a = 0.0, b = 0.0;
  if (i < g.s1) {
    const MemberDev<T> &m = md[g.member];
    const double zn = (double)gptr(m.x0)[gptr(m.iperm0)[i]];
    const double ri = gptr(m.r)[i];
    a = ri * zn;
    b = ri * gptr(m.z)[i];
    gptr(m.z)[i] = zn;
  }
  a = block_sum(a, sh);
  b = block_sum(b, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[stride + blockIdx.x] = b;
  }
}

template <class T>
__global__ __launch_bounds__(256) void kb_p_update(const Seg *__restrict__ segs, int nseg,
                                                   const MemberDev<T> *__restrict__ md) {
  const Seg g = find_seg(segs, nseg);
  const int64_t i = (int64_t)g.s0 + (int64_t)(blockIdx.x - g.block0) * 256 + threadIdx.x;
  if (i < g.s1) {
    const MemberDev<T> &m = md[g.member];
    gptr(m.p)[i] = gptr(m.z)[i] + m.sc->beta * gptr(m.p)[i];
  }
}

// copy every member's scalar block into one array (one D2H copy per step)
__global__ void kb_gather_scalars(CgScalars *const *__restrict__ sc, int nm, CgScalars *__restrict__ out) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < nm) out[m] = *sc[m];
}

} // namespace afe
