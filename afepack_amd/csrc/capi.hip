// capi.hip -- the C ABI (include/afepack_amd.h) over the host algorithms
// (host.cpp) and the gfx950 kernels (kernels.hpp).
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <map>
#include <unordered_map>
#include <memory>
#include <mutex>
#include <numeric>
#include <type_traits>
#include <string>
#include <vector>

#include <sys/resource.h>

#include "../../include/afepack_amd.h"
#include "host.hpp"
#include "kernels.hpp"
#include "batch_kernels.hpp"
#include "quadrature.hpp"

using namespace afe;

// ===========================================================================
// errors
// ===========================================================================
static thread_local std::string g_err;

static int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

struct HipFail {
  hipError_t e;
  const char *what;
};

#define HIPCHK(call)                                                           \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) throw HipFail{e_, #call};                            \
  } while (0)

#define API_BEGIN try {
#define API_END                                                                \
  }                                                                            \
  catch (const HipFail &h) {                                                   \
    return fail(AFE_E_HIP, std::string(h.what) + ": " + hipGetErrorString(h.e)); \
  }                                                                            \
  catch (const afe::Error &x) {                                                \
    return fail(x.code, x.msg);                                                \
  }                                                                            \
  catch (const std::bad_alloc &) {                                             \
    return fail(AFE_E_NOMEM, "host allocation failed");                        \
  }                                                                            \
  catch (const std::exception &x) {                                            \
    return fail(AFE_E_ARG, x.what());                                          \
  }                                                                            \
  return AFE_OK;

extern "C" const char *afe_last_error(void) { return g_err.c_str(); }
extern "C" int afe_version(void) { return 1; }

// ===========================================================================
// context and device buffers
// ===========================================================================
// chunks of a pipelined assembly (element kernels on the context stream,
// gathers on the second stream)
constexpr int kAsmChunksMax = 16;

// seconds the calling thread has waited for context locks (the server
// accounts its set-up threads' waits)
static thread_local double t_ctx_lock_wait = 0;

// the context lock: a recursive mutex whose contended acquisitions are timed
struct CtxMutex {
  std::recursive_mutex m;
  void lock() {
    if (m.try_lock()) return;
    const auto t0 = std::chrono::steady_clock::now();
    m.lock();
    t_ctx_lock_wait += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  bool try_lock() { return m.try_lock(); }
  void unlock() { m.unlock(); }
};

struct afe_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr; // the gathers of a pipelined assembly
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t asm_ev[kAsmChunksMax + 2] = {}; // pipelined assembly: chunk done, start, join
  CtxMutex mu;
  std::atomic<int64_t> bytes{0}; // device bytes held (read without the lock by the server)
  double *d_quad = nullptr; // all quadrature rules
  double *d_bary = nullptr; // reference barycentrics of every rule point (same indexing)
  int64_t lds_bytes = 0;    // LDS granted to the coarse-tail workgroup
  std::map<void *, int64_t> user_bytes; // afe_dev_alloc buffers
};

// device bytes allocated minus freed by the calling thread (the server
// measures what one request's set-up keeps)
static thread_local int64_t t_dev_bytes = 0;
// hierarchies built by the calling thread colour every level first-fit in
// index order (DSatur's fewer colours cost host time a busy set-up pool
// cannot spare; the server's set-up threads set it)
static thread_local bool t_first_fit_only = false;

// AFE_STREAM_ALLOC=1 takes the library's own device buffers from the
// device's stream-ordered pool (hipMallocAsync / hipFreeAsync on the owning
// context's stream) instead of hipMalloc / hipFree: a serving set-up
// allocates and a cache eviction frees a few hundred buffers, which the pool
// serves without a device-wide allocation or free each time.  A buffer is
// used on its context's stream, or on another stream only after an event or
// synchronisation that follows its allocation; it is freed only when no
// queued work uses it.  Off by default: it is covered by the GPU tests and
// by short serving streams, and a long serving stream run with it (and with
// further serving changes since taken out) ended in a host fault whose cause
// was not found.
static bool stream_ordered_alloc() {
  static const bool v = [] {
    const char *e = getenv("AFE_STREAM_ALLOC");
    return e && atoi(e) != 0;
  }();
  return v;
}

template <class T> struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  afe_ctx *ctx = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept { *this = std::move(o); }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) {
      release();
      p = o.p; n = o.n; ctx = o.ctx;
      o.p = nullptr; o.n = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) {
      if (ctx && stream_ordered_alloc())
        (void)hipFreeAsync(p, ctx->stream);
      else
        (void)hipFree(p);
      if (ctx) ctx->bytes -= (int64_t)(n * sizeof(T));
      t_dev_bytes -= (int64_t)(n * sizeof(T));
    }
    p = nullptr;
    n = 0;
  }
  void alloc(afe_ctx *c, size_t count) {
    if (p && n == count) return;
    release();
    ctx = c;
    if (count == 0) return;
    const hipError_t e = stream_ordered_alloc() ? hipMallocAsync((void **)&p, count * sizeof(T), c->stream)
                                                : hipMalloc((void **)&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      throw afe::Error{AFE_E_NOMEM, "device allocation of " + std::to_string(count * sizeof(T)) + " bytes failed"};
    }
    n = count;
    ctx->bytes += (int64_t)(n * sizeof(T));
    t_dev_bytes += (int64_t)(n * sizeof(T));
  }
  void upload(afe_ctx *c, const T *h, size_t count) {
    alloc(c, count);
    if (count) HIPCHK(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, c->stream));
  }
  void upload(afe_ctx *c, const std::vector<T> &v) { upload(c, v.data(), v.size()); }
};

static inline unsigned nblk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// launch tuning of the sliced-ELL family, chosen by the autotuner:
// xcd: XCD-aware logical block order; spmv_blocks: grid of the grid-stride
// SpMV / residual kernels (a multiple of 8)
struct Tuning {
  int xcd = 0;
  int spmv_blocks = 2048;
  int64_t tail_rows = 4096;          // levels at most this size run in the tail
  int64_t stream_bytes = 128 << 20;  // matrices above this use streaming loads
  uint64_t epoch = 0;                // bumped by every afe_set_tuning (captured graphs go stale)
};
static Tuning g_tune;

// level matrices larger than this many bytes use the streamed kernel forms
// (non-temporal loads, colour-coded columns); AFE_STREAM_BYTES overrides the
// threshold for hierarchies built afterwards (tests set 0 to run the
// streamed forms on small systems)
static int64_t stream_bytes() {
  const char *e = getenv("AFE_STREAM_BYTES");
  return e ? std::strtoll(e, nullptr, 10) : g_tune.stream_bytes;
}

// AFE_VERBOSE=1 prints solver progress and set-up phase times to stderr
static bool verbose() {
  static const bool v = getenv("AFE_VERBOSE") != nullptr;
  return v;
}
static constexpr int kMaxGridBlocks = 16384; // bound of spmv_blocks (partial-sum buffer)
static constexpr int64_t NAT_WINDOW = 256;    // length-sort window of the natural-order copy
static constexpr int32_t COLOUR_WINDOW = 1024; // length-sort window inside a colour

static inline unsigned sblk(int64_t nslices) {
  unsigned b = nblk(nslices, WPB);
  return g_tune.xcd == 1 ? (b + 7) / 8 * 8 : b;
}

// ---------------------------------------------------------------------------
// Scratch of afe_error_qp_noalloc: the partial sums of its reduction live in
// this file-scope device array (one copy per GPU, held until the process
// ends), so that entry point allocates nothing.  Host code never passes the
// array's name to a kernel (on the host it is only a stand-in): it uses the
// device address fetched once per device by hipGetSymbolAddress at context
// creation, with that device current.
// ---------------------------------------------------------------------------
constexpr int NOALLOC_SCRATCH_DOUBLES = RED_BLOCKS; // one partial sum per reduction block
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];
constexpr int kMaxDevices = 64;
static double *g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// with `device` current: its copy of the scratch array, looked up once
static void scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) throw afe::Error{AFE_E_NODEVICE, "device index beyond the scratch table"};
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void *p = nullptr;
  HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch_ptr[device] = (double *)p;
}

static double *scratch_of(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}

extern "C" int afe_ctx_create(int device, afe_ctx **out) {
  API_BEGIN
  if (!out) return fail(AFE_E_ARG, "null output");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return fail(AFE_E_NODEVICE, "no HIP device visible");
  }
  if (device < 0 || device >= n) return fail(AFE_E_NODEVICE, "device index out of range");
  auto *c = new afe_ctx();
  c->device = device;
  HIPCHK(hipSetDevice(device));
  scratch_init(device);
  HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HIPCHK(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
  if (stream_ordered_alloc()) {
    // the pool keeps freed memory for reuse instead of returning it to the
    // driver at every synchronisation
    hipMemPool_t pool;
    HIPCHK(hipDeviceGetDefaultMemPool(&pool, device));
    uint64_t keep = UINT64_MAX;
    HIPCHK(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
  }
  HIPCHK(hipEventCreate(&c->ev0));
  HIPCHK(hipEventCreate(&c->ev1));
  for (auto &e : c->asm_ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  {
    int lds = 0;
    HIPCHK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    // leave room for the static shared memory of the kernels
    c->lds_bytes = std::max(0, std::min(lds, 160 * 1024) - 4096);
    HIPCHK(hipFuncSetAttribute((const void *)k_tail<float>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)c->lds_bytes));
    HIPCHK(hipFuncSetAttribute((const void *)k_tail<double>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)c->lds_bytes));
    HIPCHK(hipFuncSetAttribute((const void *)kb_tail<float>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)c->lds_bytes));
    HIPCHK(hipFuncSetAttribute((const void *)kb_tail<double>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)c->lds_bytes));
    // the staged gather's row values and chunk exceed the default 64 KiB
    const void *staged[] = {
        (const void *)k_gather_staged<3, 0>, (const void *)k_gather_staged<3, 1>,
        (const void *)k_gather_staged<3, 2>, (const void *)k_gather_staged<3, 3>,
        (const void *)k_gather_staged<4, 0>, (const void *)k_gather_staged<4, 1>,
        (const void *)k_gather_staged<4, 2>, (const void *)k_gather_staged<4, 3>,
        (const void *)k_gather_staged<6, 0>, (const void *)k_gather_staged<6, 1>,
        (const void *)k_gather_staged<6, 2>, (const void *)k_gather_staged<6, 3>};
    for (const void *f : staged)
      HIPCHK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)c->lds_bytes));
  }
  HIPCHK(hipMalloc((void **)&c->d_quad, sizeof(kQuadData)));
  c->bytes += sizeof(kQuadData);
  HIPCHK(hipMemcpy(c->d_quad, kQuadData, sizeof(kQuadData), hipMemcpyHostToDevice));
  {
    // CoordTransform's barycentric coordinates of each rule point, computed
    // once here with the arithmetic the kernels would repeat per element
    std::vector<double> bary(sizeof(kQuadData) / sizeof(double), 0.0);
    for (int r = 0; r < kNumQuadRules; r++) {
      const QuadRule &q = kQuadRules[r];
      for (int l = 0; l < q.n; l++) {
        const int o = q.off + l * (q.tdim + 1);
        if (q.tdim == 2)
          ref_bary<2>(&kQuadData[o], &bary[o]);
        else
          ref_bary<3>(&kQuadData[o], &bary[o]);
      }
    }
    HIPCHK(hipMalloc((void **)&c->d_bary, sizeof(kQuadData)));
    c->bytes += sizeof(kQuadData);
    HIPCHK(hipMemcpy(c->d_bary, bary.data(), sizeof(kQuadData), hipMemcpyHostToDevice));
  }
  *out = c;
  API_END
}

extern "C" int afe_ctx_destroy(afe_ctx *c) {
  if (!c) return AFE_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  if (c->stream2) (void)hipStreamSynchronize(c->stream2);
  for (auto &kv : c->user_bytes) (void)hipFree(kv.first);
  (void)hipFree(c->d_quad);
  (void)hipFree(c->d_bary);
  (void)hipEventDestroy(c->ev0);
  (void)hipEventDestroy(c->ev1);
  for (auto &e : c->asm_ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return AFE_OK;
}

extern "C" int afe_ctx_synchronize(afe_ctx *c) {
  API_BEGIN
  if (!c) return fail(AFE_E_ARG, "null context");
  std::lock_guard<CtxMutex> lk(c->mu);
  HIPCHK(hipStreamSynchronize(c->stream));
  API_END
}

extern "C" int afe_ctx_device_bytes(afe_ctx *c, int64_t *bytes) {
  if (!c || !bytes) return fail(AFE_E_ARG, "null argument");
  std::lock_guard<CtxMutex> lk(c->mu);
  *bytes = c->bytes;
  return AFE_OK;
}

extern "C" int afe_dev_alloc(afe_ctx *c, int64_t bytes, void **d_ptr) {
  API_BEGIN
  if (!c || !d_ptr || bytes < 0) return fail(AFE_E_ARG, "bad allocation arguments");
  HIPCHK(hipSetDevice(c->device));
  *d_ptr = nullptr;
  if (bytes == 0) return AFE_OK;
  if (hipMalloc(d_ptr, (size_t)bytes) != hipSuccess) {
    (void)hipGetLastError();
    *d_ptr = nullptr;
    return fail(AFE_E_NOMEM, "hipMalloc failed");
  }
  std::lock_guard<CtxMutex> lk(c->mu);
  c->bytes += bytes;
  t_dev_bytes += bytes;
  c->user_bytes[*d_ptr] = bytes;
  API_END
}

extern "C" int afe_dev_free(afe_ctx *c, void *d_ptr) {
  API_BEGIN
  if (!c) return fail(AFE_E_ARG, "null context");
  if (!d_ptr) return AFE_OK;
  std::lock_guard<CtxMutex> lk(c->mu);
  auto it = c->user_bytes.find(d_ptr);
  if (it == c->user_bytes.end()) return fail(AFE_E_ARG, "pointer not allocated by afe_dev_alloc");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipFree(d_ptr));
  c->bytes -= it->second;
  t_dev_bytes -= it->second;
  c->user_bytes.erase(it);
  API_END
}

extern "C" int afe_dev_upload(afe_ctx *c, void *d_dst, const void *h_src, int64_t bytes) {
  API_BEGIN
  if (!c || (bytes > 0 && (!d_dst || !h_src))) return fail(AFE_E_ARG, "null argument");
  std::lock_guard<CtxMutex> lk(c->mu); // stream work: never inside another thread's capture
  HIPCHK(hipSetDevice(c->device));
  if (bytes > 0) HIPCHK(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  API_END
}

extern "C" int afe_dev_download(afe_ctx *c, void *h_dst, const void *d_src, int64_t bytes) {
  API_BEGIN
  if (!c || (bytes > 0 && (!d_src || !h_dst))) return fail(AFE_E_ARG, "null argument");
  std::lock_guard<CtxMutex> lk(c->mu); // stream work: never inside another thread's capture
  HIPCHK(hipSetDevice(c->device));
  if (bytes > 0) HIPCHK(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  API_END
}

extern "C" int afe_dev_copy(afe_ctx *c, void *d_dst, const void *d_src, int64_t bytes) {
  API_BEGIN
  if (!c || (bytes > 0 && (!d_src || !d_dst))) return fail(AFE_E_ARG, "null argument");
  std::lock_guard<CtxMutex> lk(c->mu); // stream work: never inside another thread's capture
  HIPCHK(hipSetDevice(c->device));
  if (bytes > 0) HIPCHK(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  API_END
}

extern "C" int afe_dev_zero(afe_ctx *c, void *d_ptr, int64_t bytes) {
  API_BEGIN
  if (!c || (bytes > 0 && !d_ptr)) return fail(AFE_E_ARG, "null argument");
  std::lock_guard<CtxMutex> lk(c->mu); // stream work: never inside another thread's capture
  HIPCHK(hipSetDevice(c->device));
  if (bytes > 0) HIPCHK(hipMemsetAsync(d_ptr, 0, bytes, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  API_END
}

// ===========================================================================
// meshes and fields
// ===========================================================================
extern "C" int afe_mesh_square(int32_t n, double jitter, uint64_t seed, double *coords,
                               int32_t *elem_vert, int32_t *vert_mark) {
  API_BEGIN
  if (n < 1 || !coords || !elem_vert || !vert_mark) return fail(AFE_E_ARG, "bad mesh arguments");
  if (jitter < 0 || jitter >= 0.25) return fail(AFE_E_ARG, "jitter must be in [0, 0.25)");
  mesh_square(n, jitter, seed, coords, elem_vert, vert_mark);
  API_END
}

extern "C" int afe_mesh_cube(int32_t n, double jitter, uint64_t seed, double *coords,
                             int32_t *elem_vert, int32_t *vert_mark) {
  API_BEGIN
  if (n < 1 || !coords || !elem_vert || !vert_mark) return fail(AFE_E_ARG, "bad mesh arguments");
  if (jitter < 0 || jitter >= 0.1) return fail(AFE_E_ARG, "jitter must be in [0, 0.1)");
  mesh_cube(n, jitter, seed, coords, elem_vert, vert_mark);
  API_END
}

extern "C" int afe_mesh_displace(int32_t dim, int64_t n_vert, double *coords,
                                 const int32_t *vert_mark, double amp_h, int32_t k) {
  API_BEGIN
  if ((dim != 2 && dim != 3) || !coords || !vert_mark) return fail(AFE_E_ARG, "bad arguments");
  // smooth displacement amp_h * sin(2 pi k x_d) per coordinate, interior only
  for (int64_t v = 0; v < n_vert; v++) {
    if (vert_mark[v]) continue;
    double *x = coords + v * dim;
    double d[3];
    for (int c = 0; c < dim; c++) d[c] = amp_h * std::sin(2.0 * M_PI * k * x[c]);
    for (int c = 0; c < dim; c++) x[c] += d[c];
  }
  API_END
}

extern "C" int afe_field_random(int32_t n_modes, int32_t dim, double amplitude, uint64_t seed,
                                double *k, double *phi, double *w, double *scale) {
  API_BEGIN
  if (n_modes < 1 || (dim != 2 && dim != 3) || !k || !phi || !w || !scale)
    return fail(AFE_E_ARG, "bad field arguments");
  // k_m in {1..4}^dim, phi_m in [0, 2 pi), w_m = +-(0.5 + 0.5 u); scale makes
  // sum |w_m| * scale = amplitude, so |g| <= amplitude and a = exp(g) has a
  // contrast of at most exp(2 amplitude)
  Rng rng(seed);
  double sw = 0.0;
  for (int m = 0; m < n_modes; m++) {
    for (int d = 0; d < dim; d++) k[m * dim + d] = (double)(1 + (int)(rng.uniform() * 4.0));
    phi[m] = 2.0 * M_PI * rng.uniform();
    const double mag = 0.5 + 0.5 * rng.uniform();
    w[m] = rng.uniform() < 0.5 ? -mag : mag;
    sw += mag;
  }
  *scale = amplitude / sw;
  API_END
}

// ===========================================================================
// FE space
// ===========================================================================
struct afe_space {
  afe_ctx *ctx = nullptr;
  Space h;
  int64_t seg_max = 0; // largest 256-row value segment (gather LDS)
  // per 256-row gather block: the largest element adjacent to a row of this
  // block or of an earlier one (the element chunks a block's gather waits for)
  std::vector<int64_t> gather_dep;
  bool distinct_dofs = false; // every element's DOFs are distinct (the batched gather)
  DevBuf<double> coords;
  DevBuf<int32_t> ev, elem_dof, dof_mark, adj, col;
  DevBuf<int64_t> adj_ptr, rp;
  DevBuf<uint8_t> slot;
};

// per 256-row block, the largest element adjacent to any row of blocks
// 0..b (adjacency lists are ascending, so a row's last entry is its largest)
static std::vector<int64_t> gather_deps(const Space &h) {
  const int64_t nb = (h.n_dof + 255) / 256;
  std::vector<int64_t> dep(nb);
  int64_t m = -1;
  for (int64_t b = 0; b < nb; b++) {
    const int64_t r1 = std::min<int64_t>((b + 1) * 256, h.n_dof);
    for (int64_t i = b * 256; i < r1; i++)
      if (h.adj_ptr[i + 1] > h.adj_ptr[i]) m = std::max<int64_t>(m, h.adj[h.adj_ptr[i + 1] - 1] >> 3);
    dep[b] = m;
  }
  return dep;
}

// no element lists a DOF twice (a degenerate element with a repeated vertex
// would)
static bool distinct_element_dofs(const Space &h) {
  const int nl = h.n_loc;
  for (int64_t e = 0; e < h.n_elem; e++) {
    const int32_t *d = &h.elem_dof[(size_t)e * nl];
    for (int a = 0; a < nl; a++)
      for (int b = a + 1; b < nl; b++)
        if (d[a] == d[b]) return false;
  }
  return true;
}

extern "C" int afe_space_create_sides(afe_ctx *ctx, int32_t kind, int64_t n_vert, const double *coords,
                                      int64_t n_elem, const int32_t *elem_vert, const int32_t *vert_mark,
                                      int64_t n_side, const int32_t *side_vert, const int32_t *side_mark,
                                      afe_space **out) {
  API_BEGIN
  if (!out || !coords || !elem_vert || !vert_mark || n_vert <= 0 || n_elem <= 0)
    return fail(AFE_E_ARG, "bad space arguments");
  if (kind != AFE_P1_TRI && kind != AFE_P2_TRI && kind != AFE_P1_TET)
    return fail(AFE_E_ARG, "unknown element kind");
  if (n_side < 0 || (n_side > 0 && (!side_vert || !side_mark))) return fail(AFE_E_ARG, "bad side list");
  auto s = std::make_unique<afe_space>();
  s->ctx = ctx;
  space_build(s->h, kind, n_vert, coords, n_elem, elem_vert, vert_mark, n_side, side_vert, side_mark);
  if (!ctx) { // host-only space: numbering and pattern, nothing uploaded
    *out = s.release();
    return AFE_OK;
  }
  std::lock_guard<CtxMutex> lk(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  if (s->h.rp[s->h.n_dof] > INT32_MAX) return fail(AFE_E_ARG, "nnz exceeds 2^31");
  for (int64_t r0 = 0; r0 < s->h.n_dof; r0 += 256) {
    const int64_t r1 = std::min<int64_t>(r0 + 256, s->h.n_dof);
    s->seg_max = std::max(s->seg_max, s->h.rp[r1] - s->h.rp[r0]);
  }
  s->gather_dep = gather_deps(s->h);
  s->distinct_dofs = distinct_element_dofs(s->h);
  if (s->seg_max * 8 > 64 * 1024) return fail(AFE_E_ARG, "rows too long for the LDS gather");
  s->coords.upload(ctx, s->h.coords);
  s->ev.upload(ctx, s->h.ev);
  s->elem_dof.upload(ctx, s->h.elem_dof);
  s->dof_mark.upload(ctx, s->h.dof_mark);
  s->adj.upload(ctx, s->h.adj);
  s->adj_ptr.upload(ctx, s->h.adj_ptr);
  s->slot.upload(ctx, s->h.slot);
  s->rp.upload(ctx, s->h.rp);
  s->col.upload(ctx, s->h.col);
  HIPCHK(hipStreamSynchronize(ctx->stream));
  *out = s.release();
  API_END
}

extern "C" int afe_space_create(afe_ctx *ctx, int32_t kind, int64_t n_vert, const double *coords,
                                int64_t n_elem, const int32_t *elem_vert, const int32_t *vert_mark,
                                afe_space **out) {
  return afe_space_create_sides(ctx, kind, n_vert, coords, n_elem, elem_vert, vert_mark, 0, nullptr, nullptr, out);
}

extern "C" int afe_space_destroy(afe_space *s) {
  if (!s) return AFE_OK;
  if (!s->ctx) {
    delete s;
    return AFE_OK;
  }
  std::lock_guard<CtxMutex> lk(s->ctx->mu);
  (void)hipSetDevice(s->ctx->device);
  delete s;
  return AFE_OK;
}

extern "C" int afe_space_info(const afe_space *s, int64_t *n_dof, int64_t *nnz, int32_t *n_loc) {
  if (!s) return fail(AFE_E_ARG, "null space");
  if (n_dof) *n_dof = s->h.n_dof;
  if (nnz) *nnz = s->h.rp[s->h.n_dof];
  if (n_loc) *n_loc = s->h.n_loc;
  return AFE_OK;
}

extern "C" int afe_space_dofs(const afe_space *s, int32_t *elem_dof, int32_t *dof_mark,
                              double *dof_point) {
  if (!s) return fail(AFE_E_ARG, "null space");
  const Space &h = s->h;
  if (elem_dof) std::memcpy(elem_dof, h.elem_dof.data(), h.elem_dof.size() * sizeof(int32_t));
  if (dof_mark) std::memcpy(dof_mark, h.dof_mark.data(), h.dof_mark.size() * sizeof(int32_t));
  if (dof_point) std::memcpy(dof_point, h.dof_point.data(), h.dof_point.size() * sizeof(double));
  return AFE_OK;
}

extern "C" int afe_space_pattern(const afe_space *s, int64_t *row_ptr, int64_t *col) {
  if (!s) return fail(AFE_E_ARG, "null space");
  const Space &h = s->h;
  if (row_ptr) std::memcpy(row_ptr, h.rp.data(), h.rp.size() * sizeof(int64_t));
  if (col)
    for (size_t k = 0; k < h.col.size(); k++) col[k] = h.col[k];
  return AFE_OK;
}

extern "C" int afe_space_update_coords(afe_space *s, const double *coords) {
  API_BEGIN
  if (!s || !coords) return fail(AFE_E_ARG, "null argument");
  if (!s->ctx) {
    std::memcpy(s->h.coords.data(), coords, s->h.coords.size() * sizeof(double));
    return AFE_OK;
  }
  std::lock_guard<CtxMutex> lk(s->ctx->mu);
  HIPCHK(hipSetDevice(s->ctx->device));
  std::memcpy(s->h.coords.data(), coords, s->h.coords.size() * sizeof(double));
  s->coords.upload(s->ctx, s->h.coords);
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  API_END
}

// ===========================================================================
// matrices and assembly
// ===========================================================================
struct afe_matrix {
  afe_ctx *ctx = nullptr;
  afe_space *s = nullptr; // null for a matrix given as CSR
  int64_t m = 0, n = 0, nnz = 0;
  std::vector<int64_t> h_rp; // only for a CSR-given matrix
  std::vector<int32_t> h_col;
  DevBuf<int64_t> own_rp;
  DevBuf<int32_t> own_col, own_mark;
  DevBuf<double> val;
  DevBuf<double> Ke; // element matrices of the last assembly
  int ke_layout = 0;  // how Ke holds them (k_gather's LAYOUT)
  DevBuf<int64_t> t_ptr, t_k; // transposed index map (afe_matrix_tvmult)
  DevBuf<int32_t> t_row;
  // a matrix of two FE spaces on one mesh (afe_matrix_create_pair): rows are
  // DOFs of ps0, columns DOFs of ps1; its pattern sits in h_rp/h_col and
  // own_rp/own_col, s stays null
  afe_space *ps0 = nullptr, *ps1 = nullptr;
  DevBuf<uint8_t> pair_slot; // PairPattern::slot
  int64_t pair_seg = 0;
  int pair_max_coupling = 0;
  const int64_t *rp() const { return s ? s->rp.p : own_rp.p; }
  const int32_t *col() const { return s ? s->col.p : own_col.p; }
  const std::vector<int64_t> &hrp() const { return s ? s->h.rp : h_rp; }
  const std::vector<int32_t> &hcol() const { return s ? s->h.col : h_col; }
};

extern "C" int afe_matrix_create(afe_space *s, afe_matrix **out) {
  API_BEGIN
  if (!s || !out) return fail(AFE_E_ARG, "null argument");
  if (!s->ctx) return fail(AFE_E_ARG, "host-only space (no device context)");
  std::lock_guard<CtxMutex> lk(s->ctx->mu);
  HIPCHK(hipSetDevice(s->ctx->device));
  auto a = std::make_unique<afe_matrix>();
  a->ctx = s->ctx;
  a->s = s;
  a->m = a->n = s->h.n_dof;
  a->nnz = s->h.rp[s->h.n_dof];
  a->val.alloc(s->ctx, a->nnz);
  HIPCHK(hipMemsetAsync(a->val.p, 0, a->nnz * sizeof(double), s->ctx->stream));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  *out = a.release();
  API_END
}

// a square matrix given as the reference's CSR (diagonal first in each row)
extern "C" int afe_matrix_create_csr(afe_ctx *ctx, int64_t m, const int64_t *row_ptr,
                                     const int64_t *col, const double *val, afe_matrix **out) {
  API_BEGIN
  if (!ctx || !out || !row_ptr || !col || m <= 0) return fail(AFE_E_ARG, "null argument");
  std::lock_guard<CtxMutex> lk(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  auto a = std::make_unique<afe_matrix>();
  a->ctx = ctx;
  a->m = a->n = m;
  a->nnz = row_ptr[m];
  if (a->nnz > INT32_MAX) return fail(AFE_E_ARG, "nnz exceeds 2^31");
  a->h_rp.assign(row_ptr, row_ptr + m + 1);
  a->h_col.resize(a->nnz);
  for (int64_t i = 0; i < m; i++) {
    if (row_ptr[i + 1] <= row_ptr[i] || col[row_ptr[i]] != i)
      return fail(AFE_E_ARG, "each row must start with its diagonal entry");
    for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
      if (col[k] < 0 || col[k] >= m) return fail(AFE_E_ARG, "column index out of range");
      a->h_col[k] = (int32_t)col[k];
    }
  }
  a->own_rp.upload(ctx, a->h_rp);
  a->own_col.upload(ctx, a->h_col);
  a->val.alloc(ctx, a->nnz);
  if (val)
    HIPCHK(hipMemcpyAsync(a->val.p, val, a->nnz * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  else
    HIPCHK(hipMemsetAsync(a->val.p, 0, a->nnz * sizeof(double), ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  *out = a.release();
  API_END
}

// BilinearOperator over two FE spaces of one mesh (the same-mesh branch of
// BilinearOperator.templates.h:64-172): the rectangular pattern and the
// gather map of its element matrices, values zero
extern "C" int afe_matrix_create_pair(afe_space *s0, afe_space *s1, afe_matrix **out) {
  API_BEGIN
  if (!s0 || !s1 || !out) return fail(AFE_E_ARG, "null argument");
  if (!s0->ctx || s0->ctx != s1->ctx) return fail(AFE_E_ARG, "the two spaces need one device context");
  std::lock_guard<CtxMutex> lk(s0->ctx->mu);
  HIPCHK(hipSetDevice(s0->ctx->device));
  PairPattern pp;
  pair_build(pp, s0->h, s1->h);
  if (pp.seg_max * 8 > 64 * 1024) return fail(AFE_E_ARG, "rows too long for the LDS gather");
  auto a = std::make_unique<afe_matrix>();
  a->ctx = s0->ctx;
  a->ps0 = s0;
  a->ps1 = s1;
  a->m = pp.m;
  a->n = pp.n;
  a->nnz = pp.rp[pp.m];
  a->h_rp = std::move(pp.rp);
  a->h_col = std::move(pp.col);
  a->pair_seg = pp.seg_max;
  a->pair_max_coupling = pp.max_coupling;
  a->own_rp.upload(a->ctx, a->h_rp);
  a->own_col.upload(a->ctx, a->h_col);
  a->pair_slot.upload(a->ctx, pp.slot);
  a->val.alloc(a->ctx, a->nnz);
  HIPCHK(hipMemsetAsync(a->val.p, 0, a->nnz * sizeof(double), a->ctx->stream));
  HIPCHK(hipStreamSynchronize(a->ctx->stream));
  *out = a.release();
  API_END
}

extern "C" int afe_matrix_info(const afe_matrix *a, int64_t *m, int64_t *n, int64_t *nnz, int32_t *max_coupling) {
  API_BEGIN
  if (!a) return fail(AFE_E_ARG, "null matrix");
  if (m) *m = a->m;
  if (n) *n = a->n;
  if (nnz) *nnz = a->nnz;
  if (max_coupling) *max_coupling = a->ps0 ? a->pair_max_coupling : -1;
  API_END
}

extern "C" int afe_matrix_pattern(const afe_matrix *a, int64_t *rp, int64_t *col) {
  API_BEGIN
  if (!a || !rp || !col) return fail(AFE_E_ARG, "null argument");
  const std::vector<int64_t> &r = a->hrp();
  const std::vector<int32_t> &c = a->hcol();
  for (int64_t i = 0; i <= a->m; i++) rp[i] = r[i];
  for (int64_t k = 0; k < a->nnz; k++) col[k] = c[k];
  API_END
}

extern "C" int afe_matrix_destroy(afe_matrix *a) {
  if (!a) return AFE_OK;
  std::lock_guard<CtxMutex> lk(a->ctx->mu);
  (void)hipSetDevice(a->ctx->device);
  delete a;
  return AFE_OK;
}

static FieldDev upload_field(afe_ctx *ctx, const afe_field *f, DevBuf<double> &store) {
  FieldDev fd{0, 0, nullptr, nullptr, nullptr, 0.0};
  if (!f || f->n_modes <= 0) return fd;
  const int nm = f->n_modes, dim = f->dim;
  std::vector<double> h((size_t)nm * dim + 2 * nm);
  std::memcpy(h.data(), f->k, sizeof(double) * nm * dim);
  std::memcpy(h.data() + nm * dim, f->phi, sizeof(double) * nm);
  std::memcpy(h.data() + nm * dim + nm, f->w, sizeof(double) * nm);
  store.upload(ctx, h);
  fd.n_modes = nm;
  fd.dim = dim;
  fd.k = store.p;
  fd.phi = store.p + nm * dim;
  fd.w = store.p + nm * dim + nm;
  fd.scale = f->scale;
  return fd;
}

// element matrices stored element-major on tetrahedral meshes (a row's
// elements lie far apart in index there), entry-major on triangles;
// AFE_KE_LAYOUT=entry / element forces either
static bool ke_element_major(int kind) {
  static const int forced = [] {
    const char *e = getenv("AFE_KE_LAYOUT");
    return e ? (std::strcmp(e, "element") == 0 ? 1 : 0) : -1;
  }();
  return forced >= 0 ? forced == 1 : kind == P1_TET;
}

// element matrices of the elements [e0, e1) on stream st
template <int KIND>
static void launch_elements_kind(afe_matrix *a, int op, const QuadRule *q, FieldDev fd, int64_t e0, int64_t e1,
                                 hipStream_t st) {
  afe_space *s = a->s;
  const int64_t ne = s->h.n_elem;
  const double *qd = s->ctx->d_quad + q->off;
  const double *qb = s->ctx->d_bary + q->off;
  const int nb = nblk(e1 - e0);
  const bool em = ke_element_major(KIND);
  auto go = [&](auto kern) {
    kern<<<nb, 256, 0, st>>>(ne, e0, e1, s->coords.p, s->ev.p, qd, qb, q->n, fd, a->Ke.p);
  };
  if (op == AFE_OP_STIFF)
    em ? go(k_element_matrices<KIND, 0, true>) : go(k_element_matrices<KIND, 0, false>);
  else if (op == AFE_OP_MASS)
    em ? go(k_element_matrices<KIND, 1, true>) : go(k_element_matrices<KIND, 1, false>);
  else
    em ? go(k_element_matrices<KIND, 2, true>) : go(k_element_matrices<KIND, 2, false>);
  HIPCHK(hipGetLastError());
  a->ke_layout = op == AFE_OP_MASS ? (em ? 0 : 2) : (em ? 3 : 1);
}

static void launch_elements(afe_matrix *a, int op, const QuadRule *q, FieldDev fd, int64_t e0, int64_t e1,
                            hipStream_t st) {
  switch (a->s->h.kind) {
  case P1_TRI: launch_elements_kind<P1_TRI>(a, op, q, fd, e0, e1, st); break;
  case P2_TRI: launch_elements_kind<P2_TRI>(a, op, q, fd, e0, e1, st); break;
  default: launch_elements_kind<P1_TET>(a, op, q, fd, e0, e1, st); break;
  }
}

// gather of the 256-row blocks [b0, b1) on stream st
// Row blocks go to the XCDs in contiguous ranges (xcd = 1) on tetrahedral
// meshes, whose rows have long element lists that share matrices through
// one L2 (C3 gather 2.68 -> 2.52 ms), and round-robin on triangles (C2
// 302 -> 310 us the other way); AFE_GATHER_XCD=0/1 forces either
static int gather_xcd(int kind) {
  static const int forced = [] {
    const char *e = getenv("AFE_GATHER_XCD");
    return e ? (atoi(e) != 0) : -1;
  }();
  return forced >= 0 ? forced : kind == P1_TET;
}

// The LDS-staged gather (a block's element contributions loaded together,
// then summed row by row) on tetrahedral meshes, whose rows have ~24
// elements; AFE_GATHER_STAGED=0/1 forces either form
static bool gather_staged(int kind) {
  static const int forced = [] {
    const char *e = getenv("AFE_GATHER_STAGED");
    return e ? (atoi(e) != 0) : -1;
  }();
  return forced >= 0 ? forced == 1 : kind == P1_TET;
}

template <int NL, int LAYOUT> static void launch_gather_nl(afe_matrix *a, int64_t b0, int64_t b1, hipStream_t st) {
  afe_space *s = a->s;
  const size_t lds = (size_t)s->seg_max * sizeof(double);
  const int xcd = gather_xcd(s->h.kind);
  const int64_t nb = xcd ? (b1 - b0 + 7) / 8 * 8 : b1 - b0;
  const int64_t seg_cap = (s->seg_max + 1) / 2 * 2; // keeps the staged values 16-byte aligned
  const size_t lds_staged = (size_t)seg_cap * sizeof(double) + (size_t)GATHER_CHUNK * NL * (sizeof(double) + 1);
  if (s->distinct_dofs && gather_staged(s->h.kind) && (int64_t)lds_staged <= a->ctx->lds_bytes) {
    k_gather_staged<NL, LAYOUT><<<(unsigned)nb, 256, lds_staged, st>>>(s->h.n_dof, b0, b1, xcd, seg_cap, s->rp.p,
                                                                       s->adj_ptr.p, s->adj.p, s->slot.p,
                                                                       s->h.n_elem, a->Ke.p, a->val.p);
  } else if (s->distinct_dofs)
    k_gather<NL, LAYOUT, true><<<(unsigned)nb, 256, lds, st>>>(s->h.n_dof, b0, b1, xcd, s->rp.p, s->adj_ptr.p,
                                                               s->adj.p, s->slot.p, s->h.n_elem, a->Ke.p, a->val.p);
  else
    k_gather<NL, LAYOUT, false><<<(unsigned)nb, 256, lds, st>>>(s->h.n_dof, b0, b1, xcd, s->rp.p, s->adj_ptr.p,
                                                                s->adj.p, s->slot.p, s->h.n_elem, a->Ke.p, a->val.p);
}

template <int NL> static void launch_gather_layout(afe_matrix *a, int64_t b0, int64_t b1, hipStream_t st) {
  switch (a->ke_layout) {
  case 0: launch_gather_nl<NL, 0>(a, b0, b1, st); break;
  case 1: launch_gather_nl<NL, 1>(a, b0, b1, st); break;
  case 2: launch_gather_nl<NL, 2>(a, b0, b1, st); break;
  default: launch_gather_nl<NL, 3>(a, b0, b1, st); break;
  }
}

static int64_t gather_blocks(const afe_space *s) { return (s->h.n_dof + 255) / 256; }

static void launch_gather(afe_matrix *a, int64_t b0, int64_t b1, hipStream_t st) {
  if (b1 <= b0) return;
  switch (a->s->h.n_loc) {
  case 3: launch_gather_layout<3>(a, b0, b1, st); break;
  case 4: launch_gather_layout<4>(a, b0, b1, st); break;
  default: launch_gather_layout<6>(a, b0, b1, st); break;
  }
  HIPCHK(hipGetLastError());
}

static void launch_gather(afe_matrix *a) { launch_gather(a, 0, gather_blocks(a->s), a->ctx->stream); }

// number of element chunks of a pipelined assembly: 8 for tetrahedral
// meshes from 2^18 elements up, where the gather (rows of ~24 elements) is
// latency bound and leaves the chip room for the next chunk's element
// matrices (C3, variable coefficient: 5.3 -> 4.0 ms; constant: no change);
// 1 (no pipeline) on triangles, where the gather streams at bandwidth and
// the overlap measured slower (C2 0.50 -> 0.53 ms, P2 variable coefficient
// 1.37 -> 1.63 ms).  AFE_ASM_CHUNKS=k forces k chunks.
static int asm_chunks(int kind, int64_t n_elem) {
  static const int forced = [] {
    const char *e = getenv("AFE_ASM_CHUNKS");
    return e ? std::max(1, std::min(kAsmChunksMax, atoi(e))) : 0;
  }();
  if (forced) return forced;
  return kind == P1_TET && n_elem >= (int64_t(1) << 18) ? 8 : 1;
}

// Element matrices and CSR gather of afe_assemble.  Pipelined: the elements
// are computed in chunks on the context stream; after chunk k the gather of
// every row block whose adjacent elements all lie in chunks 0..k runs on the
// second stream, so the element work of the next chunk overlaps the gather
// of the rows already complete.  Each row is still summed by one thread in
// ascending element order, so the values are the same bits as the two
// whole-mesh launches.  The context stream waits for the last gather.
static void launch_assembly(afe_matrix *a, int op, const QuadRule *q, FieldDev fd) {
  afe_space *s = a->s;
  afe_ctx *c = a->ctx;
  const int64_t ne = s->h.n_elem, nb = gather_blocks(s);
  const int K = (int64_t)s->gather_dep.size() == nb ? asm_chunks(s->h.kind, ne) : 1;
  if (K <= 1) {
    launch_elements(a, op, q, fd, 0, ne, c->stream);
    launch_gather(a, 0, nb, c->stream);
    return;
  }
  hipEvent_t ev_start = c->asm_ev[kAsmChunksMax], ev_join = c->asm_ev[kAsmChunksMax + 1];
  HIPCHK(hipEventRecord(ev_start, c->stream)); // the gathers follow earlier work of the context stream
  HIPCHK(hipStreamWaitEvent(c->stream2, ev_start, 0));
  const int64_t ce = (ne + K - 1) / K;
  int64_t b0 = 0;
  for (int k = 0; k < K; k++) {
    const int64_t e0 = k * ce, e1 = std::min(ne, e0 + ce);
    if (e0 >= ne) break;
    launch_elements(a, op, q, fd, e0, e1, c->stream);
    // blocks whose largest adjacent element (prefix maximum) is below e1
    const int64_t b1 = e1 == ne ? nb
                                : (int64_t)(std::lower_bound(s->gather_dep.begin(), s->gather_dep.end(), e1) -
                                            s->gather_dep.begin());
    if (b1 > b0) {
      HIPCHK(hipEventRecord(c->asm_ev[k], c->stream));
      HIPCHK(hipStreamWaitEvent(c->stream2, c->asm_ev[k], 0));
      launch_gather(a, b0, b1, c->stream2);
      b0 = b1;
    }
  }
  HIPCHK(hipEventRecord(ev_join, c->stream2));
  HIPCHK(hipStreamWaitEvent(c->stream, ev_join, 0));
}

// the two-space element matrices (entry-major, LAYOUT 2) or the caller's
// (element-major n0 x n1, LAYOUT 0) gathered into the rectangular CSR values
template <int N0, int N1, int LAYOUT> static void launch_gather_pair_nl(afe_matrix *a) {
  afe_space *s0 = a->ps0;
  const size_t lds = (size_t)a->pair_seg * sizeof(double);
  k_gather_pair<N0, N1, LAYOUT><<<nblk(a->m), 256, lds, a->ctx->stream>>>(
      a->m, a->own_rp.p, s0->adj_ptr.p, s0->adj.p, a->pair_slot.p, s0->h.n_elem, a->Ke.p, a->val.p);
}

template <int LAYOUT> static void launch_gather_pair(afe_matrix *a) {
  const int n0 = a->ps0->h.n_loc, n1 = a->ps1->h.n_loc;
  if (n0 == 3 && n1 == 3) launch_gather_pair_nl<3, 3, LAYOUT>(a);
  else if (n0 == 3 && n1 == 6) launch_gather_pair_nl<3, 6, LAYOUT>(a);
  else if (n0 == 6 && n1 == 3) launch_gather_pair_nl<6, 3, LAYOUT>(a);
  else if (n0 == 6 && n1 == 6) launch_gather_pair_nl<6, 6, LAYOUT>(a);
  else launch_gather_pair_nl<4, 4, LAYOUT>(a);
  HIPCHK(hipGetLastError());
}

template <int K0, int K1> static void launch_mass_pair(afe_matrix *a, const QuadRule *q) {
  afe_space *s0 = a->ps0;
  const int64_t ne = s0->h.n_elem;
  k_element_mass_pair<K0, K1><<<nblk(ne), 256, 0, a->ctx->stream>>>(
      ne, s0->coords.p, s0->ev.p, s0->ctx->d_quad + q->off, s0->ctx->d_bary + q->off, q->n, a->Ke.p);
  HIPCHK(hipGetLastError());
}

// L2InnerProduct of two spaces on one mesh (BilinearOperator.templates.h:
// 300-326, the quadrature of element 0)
static int assemble_pair(afe_matrix *a, int32_t op, int32_t acc) {
  if (op != AFE_OP_MASS) return fail(AFE_E_ARG, "two-space matrices assemble the L2 inner product (AFE_OP_MASS) only");
  afe_space *s0 = a->ps0;
  const QuadRule *q = quad_find(s0->h.tdim, acc);
  if (!q) return fail(AFE_E_ARG, "no quadrature rule of that accuracy");
  std::lock_guard<CtxMutex> lk(a->ctx->mu);
  HIPCHK(hipSetDevice(a->ctx->device));
  a->Ke.alloc(a->ctx, (size_t)s0->h.n_elem * s0->h.n_loc * a->ps1->h.n_loc);
  const int k0 = s0->h.kind, k1 = a->ps1->h.kind;
  if (k0 == P1_TRI && k1 == P1_TRI) launch_mass_pair<P1_TRI, P1_TRI>(a, q);
  else if (k0 == P1_TRI && k1 == P2_TRI) launch_mass_pair<P1_TRI, P2_TRI>(a, q);
  else if (k0 == P2_TRI && k1 == P1_TRI) launch_mass_pair<P2_TRI, P1_TRI>(a, q);
  else if (k0 == P2_TRI && k1 == P2_TRI) launch_mass_pair<P2_TRI, P2_TRI>(a, q);
  else launch_mass_pair<P1_TET, P1_TET>(a, q);
  launch_gather_pair<2>(a);
  HIPCHK(hipStreamSynchronize(a->ctx->stream));
  return AFE_OK;
}

extern "C" int afe_assemble(afe_matrix *a, int32_t op, int32_t acc, const afe_field *coef) {
  API_BEGIN
  if (a && a->ps0) return assemble_pair(a, op, acc);
  if (!a || !a->s) return fail(AFE_E_ARG, "matrix has no FE space");
  if (op < AFE_OP_STIFF || op > AFE_OP_VARSTIFF) return fail(AFE_E_ARG, "unknown operator");
  if (op == AFE_OP_VARSTIFF && (!coef || coef->n_modes <= 0 || coef->dim != a->s->h.tdim))
    return fail(AFE_E_ARG, "variable coefficient needs a field of the space dimension");
  afe_space *s = a->s;
  const QuadRule *q = quad_find(s->h.tdim, acc);
  if (!q) return fail(AFE_E_ARG, "no quadrature rule of that accuracy");
  std::lock_guard<CtxMutex> lk(a->ctx->mu);
  HIPCHK(hipSetDevice(a->ctx->device));
  DevBuf<double> fstore;
  FieldDev fd = upload_field(a->ctx, op == AFE_OP_VARSTIFF ? coef : nullptr, fstore);
  a->Ke.alloc(a->ctx, (size_t)s->h.n_elem * s->h.n_loc * s->h.n_loc);
  launch_assembly(a, op, q, fd);
  HIPCHK(hipStreamSynchronize(a->ctx->stream));
  API_END
}

// average durations of the two assembly kernels of afe_assemble (element
// matrices, CSR gather) over `reps` assemblies, from HIP events on the
// context stream, and the assembly's algorithmic bytes: vertex coordinates,
// element vertex and DOF indices read, CSR values written
extern "C" int afe_assemble_kernel_timing(afe_matrix *a, int32_t op, int32_t acc, const afe_field *coef,
                                          int32_t reps, double *elem_ms, double *gather_ms, double *bytes) {
  API_BEGIN
  if (a && a->ps0) return fail(AFE_E_ARG, "kernel timing covers one-space assembly only");
  if (!a || !a->s || reps < 1) return fail(AFE_E_ARG, "bad arguments");
  if (op < AFE_OP_STIFF || op > AFE_OP_VARSTIFF) return fail(AFE_E_ARG, "unknown operator");
  if (op == AFE_OP_VARSTIFF && (!coef || coef->n_modes <= 0 || coef->dim != a->s->h.tdim))
    return fail(AFE_E_ARG, "variable coefficient needs a field of the space dimension");
  afe_space *s = a->s;
  const QuadRule *q = quad_find(s->h.tdim, acc);
  if (!q) return fail(AFE_E_ARG, "no quadrature rule of that accuracy");
  std::lock_guard<CtxMutex> lk(a->ctx->mu);
  HIPCHK(hipSetDevice(a->ctx->device));
  DevBuf<double> fstore;
  FieldDev fd = upload_field(a->ctx, op == AFE_OP_VARSTIFF ? coef : nullptr, fstore);
  a->Ke.alloc(a->ctx, (size_t)s->h.n_elem * s->h.n_loc * s->h.n_loc);
  hipStream_t st = a->ctx->stream;
  hipEvent_t ev[3];
  for (auto &e : ev) HIPCHK(hipEventCreate(&e));
  double te = 0, tg = 0;
  for (int r = 0; r < reps; r++) {
    HIPCHK(hipEventRecord(ev[0], st));
    launch_elements(a, op, q, fd, 0, s->h.n_elem, st);
    HIPCHK(hipEventRecord(ev[1], st));
    launch_gather(a);
    HIPCHK(hipEventRecord(ev[2], st));
    HIPCHK(hipEventSynchronize(ev[2]));
    float m0 = 0, m1 = 0;
    HIPCHK(hipEventElapsedTime(&m0, ev[0], ev[1]));
    HIPCHK(hipEventElapsedTime(&m1, ev[1], ev[2]));
    te += m0;
    tg += m1;
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  if (elem_ms) *elem_ms = te / reps;
  if (gather_ms) *gather_ms = tg / reps;
  if (bytes) {
    const Space &h = s->h;
    const int nv = h.kind == P1_TET ? 4 : 3;
    *bytes = (double)h.coords.size() * 8.0 + (double)h.n_elem * (nv + h.n_loc) * 4.0 +
             (double)h.rp[h.n_dof] * 8.0;
  }
  API_END
}

// average device time of afe_assemble's kernels (element matrices and CSR
// gather, pipelined as afe_assemble runs them) over `reps` assemblies, from
// HIP events on the context stream, which waits for the last gather
extern "C" int afe_assemble_timing(afe_matrix *a, int32_t op, int32_t acc, const afe_field *coef, int32_t reps,
                                   double *ms) {
  API_BEGIN
  if (a && a->ps0) return fail(AFE_E_ARG, "assembly timing covers one-space assembly only");
  if (!a || !a->s || reps < 1 || !ms) return fail(AFE_E_ARG, "bad arguments");
  if (op < AFE_OP_STIFF || op > AFE_OP_VARSTIFF) return fail(AFE_E_ARG, "unknown operator");
  if (op == AFE_OP_VARSTIFF && (!coef || coef->n_modes <= 0 || coef->dim != a->s->h.tdim))
    return fail(AFE_E_ARG, "variable coefficient needs a field of the space dimension");
  afe_space *s = a->s;
  const QuadRule *q = quad_find(s->h.tdim, acc);
  if (!q) return fail(AFE_E_ARG, "no quadrature rule of that accuracy");
  std::lock_guard<CtxMutex> lk(a->ctx->mu);
  HIPCHK(hipSetDevice(a->ctx->device));
  DevBuf<double> fstore;
  FieldDev fd = upload_field(a->ctx, op == AFE_OP_VARSTIFF ? coef : nullptr, fstore);
  a->Ke.alloc(a->ctx, (size_t)s->h.n_elem * s->h.n_loc * s->h.n_loc);
  hipStream_t st = a->ctx->stream;
  hipEvent_t ev[2];
  for (auto &e : ev) HIPCHK(hipEventCreate(&e));
  double t = 0;
  for (int r = 0; r < reps; r++) {
    HIPCHK(hipEventRecord(ev[0], st));
    launch_assembly(a, op, q, fd);
    HIPCHK(hipEventRecord(ev[1], st));
    HIPCHK(hipEventSynchronize(ev[1]));
    float m = 0;
    HIPCHK(hipEventElapsedTime(&m, ev[0], ev[1]));
    t += m;
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  *ms = t / reps;
  API_END
}

extern "C" int afe_assemble_elements(afe_matrix *a, const double *h_elem_mat) {
  API_BEGIN
  if (a && a->ps0 && h_elem_mat) {
    std::lock_guard<CtxMutex> lk(a->ctx->mu);
    HIPCHK(hipSetDevice(a->ctx->device));
    a->Ke.upload(a->ctx, h_elem_mat, (size_t)a->ps0->h.n_elem * a->ps0->h.n_loc * a->ps1->h.n_loc);
    launch_gather_pair<0>(a);
    HIPCHK(hipStreamSynchronize(a->ctx->stream));
    return AFE_OK;
  }
  if (!a || !a->s || !h_elem_mat) return fail(AFE_E_ARG, "null argument");
  std::lock_guard<CtxMutex> lk(a->ctx->mu);
  HIPCHK(hipSetDevice(a->ctx->device));
  const Space &h = a->s->h;
  a->Ke.upload(a->ctx, h_elem_mat, (size_t)h.n_elem * h.n_loc * h.n_loc);
  a->ke_layout = 0;
  launch_gather(a);
  HIPCHK(hipStreamSynchronize(a->ctx->stream));
  API_END
}

extern "C" int afe_matrix_values(const afe_matrix *a, double *h_val) {
  API_BEGIN
  if (!a || !h_val) return fail(AFE_E_ARG, "null argument");
  std::lock_guard<CtxMutex> lk(a->ctx->mu);
  HIPCHK(hipSetDevice(a->ctx->device));
  HIPCHK(hipMemcpyAsync(h_val, a->val.p, a->nnz * sizeof(double), hipMemcpyDeviceToHost, a->ctx->stream));
  HIPCHK(hipStreamSynchronize(a->ctx->stream));
  API_END
}

extern "C" int afe_matrix_set_values(afe_matrix *a, const double *h_val) {
  API_BEGIN
  if (!a || !h_val) return fail(AFE_E_ARG, "null argument");
  std::lock_guard<CtxMutex> lk(a->ctx->mu);
  HIPCHK(hipSetDevice(a->ctx->device));
  HIPCHK(hipMemcpyAsync(a->val.p, h_val, a->nnz * sizeof(double), hipMemcpyHostToDevice, a->ctx->stream));
  HIPCHK(hipStreamSynchronize(a->ctx->stream));
  API_END
}

extern "C" double *afe_matrix_device_values(afe_matrix *a) { return a ? a->val.p : nullptr; }

// SparseMatrix::vmult / Tvmult (SparseMatrix.h:100-132): products summed in
// the reference's order without contraction, so the results are the
// reference's bit for bit.  Tvmult's scatter y[col[k]] += v[k] x[i] over rows
// i in ascending order is a gather here: column j sums its entries in
// ascending row order through the transposed index map (tptr, trow, tk).
#pragma clang fp contract(off)
__global__ __launch_bounds__(256) void k_csr_vmult(int64_t m, const int64_t *__restrict__ rp,
                                                   const int32_t *__restrict__ col,
                                                   const double *__restrict__ val,
                                                   const double *__restrict__ x, double *__restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double s = 0.0;
  for (int64_t k = rp[i]; k < rp[i + 1]; k++) s += val[k] * x[col[k]];
  y[i] = s;
}

__global__ __launch_bounds__(256) void k_csr_tvmult(int64_t n, const int64_t *__restrict__ tptr,
                                                    const int32_t *__restrict__ trow,
                                                    const int64_t *__restrict__ tk,
                                                    const double *__restrict__ val,
                                                    const double *__restrict__ x, double *__restrict__ y) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double s = 0.0;
  for (int64_t t = tptr[j]; t < tptr[j + 1]; t++) s += val[tk[t]] * x[trow[t]];
  y[j] = s;
}
#pragma clang fp contract(on)

extern "C" int afe_matrix_tvmult(afe_matrix *a, const double *d_x, double *d_y) {
  API_BEGIN
  if (!a || !d_x || !d_y) return fail(AFE_E_ARG, "null argument");
  std::lock_guard<CtxMutex> lk(a->ctx->mu);
  HIPCHK(hipSetDevice(a->ctx->device));
  if (!a->t_ptr.p) {
    // transposed index map of the pattern, built once per matrix: column j
    // lists (row, CSR position) in ascending row order
    const std::vector<int64_t> &rp = a->hrp();
    const std::vector<int32_t> &col = a->hcol();
    std::vector<int64_t> tptr((size_t)a->n + 1, 0), tk((size_t)a->nnz);
    std::vector<int32_t> trow((size_t)a->nnz);
    for (int64_t k = 0; k < a->nnz; k++) tptr[(size_t)col[k] + 1]++;
    for (int64_t j = 0; j < a->n; j++) tptr[j + 1] += tptr[j];
    std::vector<int64_t> fill(tptr.begin(), tptr.end() - 1);
    for (int64_t i = 0; i < a->m; i++)
      for (int64_t k = rp[i]; k < rp[i + 1]; k++) {
        const int64_t t = fill[col[k]]++;
        trow[t] = (int32_t)i;
        tk[t] = k;
      }
    a->t_ptr.upload(a->ctx, tptr);
    a->t_row.upload(a->ctx, trow);
    a->t_k.upload(a->ctx, tk);
  }
  k_csr_tvmult<<<nblk(a->n), 256, 0, a->ctx->stream>>>(a->n, a->t_ptr.p, a->t_row.p, a->t_k.p, a->val.p, d_x, d_y);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(a->ctx->stream));
  API_END
}

extern "C" int afe_matrix_vmult(afe_matrix *a, const double *d_x, double *d_y) {
  API_BEGIN
  if (!a || !d_x || !d_y) return fail(AFE_E_ARG, "null argument");
  std::lock_guard<CtxMutex> lk(a->ctx->mu);
  HIPCHK(hipSetDevice(a->ctx->device));
  k_csr_vmult<<<nblk(a->m), 256, 0, a->ctx->stream>>>(a->m, a->rp(), a->col(), a->val.p, d_x, d_y);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(a->ctx->stream));
  API_END
}

// ===========================================================================
// right-hand side, Dirichlet elimination, error functionals
// ===========================================================================
template <int KIND>
static void launch_loads(afe_space *s, const QuadRule *q, int fkind, FieldDev fd, const double *nodal,
                         double *Fe) {
  const int64_t ne = s->h.n_elem;
  k_element_loads<KIND><<<nblk(ne), 256, 0, s->ctx->stream>>>(ne, s->coords.p, s->ev.p, s->elem_dof.p,
                                                               s->ctx->d_quad + q->off, q->n, fkind, fd,
                                                               nodal, Fe);
  HIPCHK(hipGetLastError());
}

extern "C" int afe_rhs(afe_space *s, int32_t acc, int32_t fkind, const afe_field *coef,
                       const double *d_nodal, double *d_f) {
  API_BEGIN
  if (!s || !d_f) return fail(AFE_E_ARG, "null argument");
  if (!s->ctx) return fail(AFE_E_ARG, "host-only space (no device context)");
  if (fkind < AFE_F_POISSON || fkind > AFE_F_ONE) return fail(AFE_E_ARG, "unknown rhs kind");
  if (fkind == AFE_F_NODAL && !d_nodal) return fail(AFE_E_ARG, "nodal rhs needs DOF values");
  if (fkind == AFE_F_VARCOEF && (!coef || coef->dim != s->h.tdim))
    return fail(AFE_E_ARG, "variable-coefficient rhs needs a field");
  const QuadRule *q = quad_find(s->h.tdim, acc);
  if (!q) return fail(AFE_E_ARG, "no quadrature rule of that accuracy");
  std::lock_guard<CtxMutex> lk(s->ctx->mu);
  HIPCHK(hipSetDevice(s->ctx->device));
  DevBuf<double> fstore, Fe;
  FieldDev fd = upload_field(s->ctx, fkind == AFE_F_VARCOEF ? coef : nullptr, fstore);
  Fe.alloc(s->ctx, (size_t)s->h.n_elem * s->h.n_loc);
  switch (s->h.kind) {
  case P1_TRI: launch_loads<P1_TRI>(s, q, fkind, fd, d_nodal, Fe.p); break;
  case P2_TRI: launch_loads<P2_TRI>(s, q, fkind, fd, d_nodal, Fe.p); break;
  default: launch_loads<P1_TET>(s, q, fkind, fd, d_nodal, Fe.p); break;
  }
  k_gather_vec<<<nblk(s->h.n_dof), 256, 0, s->ctx->stream>>>(s->h.n_dof, s->h.n_loc, s->adj_ptr.p, s->adj.p,
                                                              Fe.p, d_f);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  API_END
}

extern "C" int afe_apply_dirichlet(afe_matrix *a, double *d_u, double *d_f) {
  API_BEGIN
  (void)d_u;
  if (a && a->ps0) return fail(AFE_E_ARG, "Dirichlet elimination needs the square matrix of one space");
  if (!a || !a->s || !d_f) return fail(AFE_E_ARG, "matrix without FE space or null rhs");
  std::lock_guard<CtxMutex> lk(a->ctx->mu);
  HIPCHK(hipSetDevice(a->ctx->device));
  afe_space *s = a->s;
  k_dirichlet_zero<<<nblk(s->h.n_dof), 256, 0, a->ctx->stream>>>(s->h.n_dof, s->rp.p, s->col.p, s->dof_mark.p,
                                                                  a->val.p, d_f);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(a->ctx->stream));
  API_END
}

extern "C" int afe_space_qpoints(afe_space *s, int32_t acc, int64_t *n_qp, double *h_pts, double *h_jxw) {
  API_BEGIN
  if (!s || !s->ctx) return fail(AFE_E_ARG, "space without device context");
  const QuadRule *q = quad_find(s->h.tdim, acc);
  if (!q) return fail(AFE_E_ARG, "no quadrature rule of that accuracy");
  if (n_qp) *n_qp = q->n;
  if (!h_pts && !h_jxw) return AFE_OK;
  std::lock_guard<CtxMutex> lk(s->ctx->mu);
  HIPCHK(hipSetDevice(s->ctx->device));
  const int64_t ne = s->h.n_elem;
  const int D = s->h.tdim;
  DevBuf<double> pts, jxw;
  pts.alloc(s->ctx, (size_t)ne * q->n * D);
  jxw.alloc(s->ctx, (size_t)ne * q->n);
  const double *qd = s->ctx->d_quad + q->off;
  hipStream_t st = s->ctx->stream;
  switch (s->h.kind) {
  case P1_TRI: k_qpoints<P1_TRI><<<nblk(ne), 256, 0, st>>>(ne, s->coords.p, s->ev.p, qd, q->n, pts.p, jxw.p); break;
  case P2_TRI: k_qpoints<P2_TRI><<<nblk(ne), 256, 0, st>>>(ne, s->coords.p, s->ev.p, qd, q->n, pts.p, jxw.p); break;
  default: k_qpoints<P1_TET><<<nblk(ne), 256, 0, st>>>(ne, s->coords.p, s->ev.p, qd, q->n, pts.p, jxw.p); break;
  }
  HIPCHK(hipGetLastError());
  if (h_pts) HIPCHK(hipMemcpyAsync(h_pts, pts.p, pts.n * sizeof(double), hipMemcpyDeviceToHost, st));
  if (h_jxw) HIPCHK(hipMemcpyAsync(h_jxw, jxw.p, jxw.n * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  API_END
}

extern "C" int afe_rhs_qp(afe_space *s, int32_t acc, const double *h_fq, double *d_f) {
  API_BEGIN
  if (!s || !s->ctx || !h_fq || !d_f) return fail(AFE_E_ARG, "null argument");
  const QuadRule *q = quad_find(s->h.tdim, acc);
  if (!q) return fail(AFE_E_ARG, "no quadrature rule of that accuracy");
  std::lock_guard<CtxMutex> lk(s->ctx->mu);
  HIPCHK(hipSetDevice(s->ctx->device));
  const int64_t ne = s->h.n_elem;
  DevBuf<double> fq, Fe;
  fq.upload(s->ctx, h_fq, (size_t)ne * q->n);
  Fe.alloc(s->ctx, (size_t)ne * s->h.n_loc);
  const double *qd = s->ctx->d_quad + q->off;
  hipStream_t st = s->ctx->stream;
  switch (s->h.kind) {
  case P1_TRI: k_element_loads_qp<P1_TRI><<<nblk(ne), 256, 0, st>>>(ne, s->coords.p, s->ev.p, qd, q->n, fq.p, Fe.p); break;
  case P2_TRI: k_element_loads_qp<P2_TRI><<<nblk(ne), 256, 0, st>>>(ne, s->coords.p, s->ev.p, qd, q->n, fq.p, Fe.p); break;
  default: k_element_loads_qp<P1_TET><<<nblk(ne), 256, 0, st>>>(ne, s->coords.p, s->ev.p, qd, q->n, fq.p, Fe.p); break;
  }
  k_gather_vec<<<nblk(s->h.n_dof), 256, 0, st>>>(s->h.n_dof, s->h.n_loc, s->adj_ptr.p, s->adj.p, Fe.p, d_f);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  API_END
}

extern "C" int afe_apply_dirichlet_values(afe_matrix *a, const int32_t *marks, int32_t n_marks, const double *h_g,
                                          double *d_u, double *d_f) {
  return afe_apply_dirichlet_bc(a, marks, n_marks, h_g, d_u, d_f, 1);
}

extern "C" int afe_apply_dirichlet_bc(afe_matrix *a, const int32_t *marks, int32_t n_marks, const double *h_g,
                                      double *d_u, double *d_f, int32_t preserve_symmetry) {
  API_BEGIN
  (void)d_u;
  if (a && a->ps0) return fail(AFE_E_ARG, "Dirichlet elimination needs the square matrix of one space");
  if (!a || !a->s || !d_f || !h_g || (n_marks > 0 && !marks)) return fail(AFE_E_ARG, "bad arguments");
  std::lock_guard<CtxMutex> lk(a->ctx->mu);
  HIPCHK(hipSetDevice(a->ctx->device));
  afe_space *s = a->s;
  const int64_t n = s->h.n_dof;
  std::vector<uint8_t> flag(n, 0);
  for (int64_t i = 0; i < n; i++) {
    const int32_t m = s->h.dof_mark[i];
    for (int32_t k = 0; k < n_marks; k++)
      if (marks[k] == m) flag[i] = 1;
  }
  DevBuf<uint8_t> dflag;
  DevBuf<double> dg, diag0;
  dflag.upload(a->ctx, flag);
  dg.upload(a->ctx, h_g, n);
  diag0.alloc(a->ctx, n);
  hipStream_t st = a->ctx->stream;
  k_csr_diag<<<nblk(n), 256, 0, st>>>(n, s->rp.p, a->val.p, diag0.p);
  k_dirichlet_values<<<nblk(n), 256, 0, st>>>(n, s->rp.p, s->col.p, dflag.p, diag0.p, dg.p, a->val.p, d_f,
                                              preserve_symmetry ? 1 : 0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  API_END
}

// the error functional's kernels: per-block partial sums of the squared
// errors into part (RED_BLOCKS doubles), then sqrt of their sum into d_err
static void launch_error_qp(afe_space *s, const QuadRule *q, int h1, const double *d_uh, const double *d_exact,
                            double *part, double *d_err) {
  const int64_t ne = s->h.n_elem;
  const double *qd = s->ctx->d_quad + q->off;
  hipStream_t st = s->ctx->stream;
  switch (s->h.kind) {
  case P1_TRI:
    k_error_qp_part<P1_TRI><<<RED_BLOCKS, 256, 0, st>>>(ne, s->coords.p, s->ev.p, s->elem_dof.p, qd, q->n, h1, d_uh,
                                                        d_exact, part);
    break;
  case P2_TRI:
    k_error_qp_part<P2_TRI><<<RED_BLOCKS, 256, 0, st>>>(ne, s->coords.p, s->ev.p, s->elem_dof.p, qd, q->n, h1, d_uh,
                                                        d_exact, part);
    break;
  default:
    k_error_qp_part<P1_TET><<<RED_BLOCKS, 256, 0, st>>>(ne, s->coords.p, s->ev.p, s->elem_dof.p, qd, q->n, h1, d_uh,
                                                        d_exact, part);
    break;
  }
  k_finalize_sqrt<<<1, FIN_THREADS, 0, st>>>(part, RED_BLOCKS, d_err);
  HIPCHK(hipGetLastError());
}

extern "C" int afe_error_qp(afe_space *s, const double *d_uh, int32_t acc, int32_t h1, const double *h_exact,
                            double *err) {
  API_BEGIN
  if (!s || !s->ctx || !d_uh || !h_exact || !err) return fail(AFE_E_ARG, "null argument");
  const QuadRule *q = quad_find(s->h.tdim, acc);
  if (!q) return fail(AFE_E_ARG, "no quadrature rule of that accuracy");
  std::lock_guard<CtxMutex> lk(s->ctx->mu);
  HIPCHK(hipSetDevice(s->ctx->device));
  const int64_t ne = s->h.n_elem;
  DevBuf<double> ex, part, tot;
  ex.upload(s->ctx, h_exact, (size_t)ne * q->n * (h1 ? s->h.tdim : 1));
  part.alloc(s->ctx, RED_BLOCKS);
  tot.alloc(s->ctx, 1);
  launch_error_qp(s, q, h1, d_uh, ex.p, part.p, tot.p);
  HIPCHK(hipMemcpyAsync(err, tot.p, sizeof(double), hipMemcpyDeviceToHost, s->ctx->stream));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  API_END
}

// afe_error_qp with the exact values and the result on the device and the
// reduction's partial sums in the file-scope scratch array: no allocation
extern "C" int afe_error_qp_noalloc(afe_space *s, const double *d_uh, int32_t acc, int32_t h1,
                                    const double *d_exact, double *d_err) {
  API_BEGIN
  if (!s || !s->ctx || !d_uh || !d_exact || !d_err) return fail(AFE_E_ARG, "null argument");
  const QuadRule *q = quad_find(s->h.tdim, acc);
  if (!q) return fail(AFE_E_ARG, "no quadrature rule of that accuracy");
  static_assert(RED_BLOCKS <= NOALLOC_SCRATCH_DOUBLES, "the reduction's partial sums must fit the scratch array");
  std::lock_guard<CtxMutex> lk(s->ctx->mu);
  HIPCHK(hipSetDevice(s->ctx->device));
  double *scr = scratch_of(s->ctx->device);
  if (!scr) return fail(AFE_E_HIP, "scratch array of this device not initialised (no context created on it)");
  // every block overwrites its partial sum: nothing to clear between calls
  launch_error_qp(s, q, h1, d_uh, d_exact, scr, d_err);
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  API_END
}

template <int KIND>
static void launch_error(afe_space *s, const QuadRule *q, int h1, const double *uh, double *out) {
  const int64_t ne = s->h.n_elem;
  k_element_error<KIND><<<nblk(ne), 256, 0, s->ctx->stream>>>(ne, s->coords.p, s->ev.p, s->elem_dof.p,
                                                               s->ctx->d_quad + q->off, q->n, h1, uh, out);
  HIPCHK(hipGetLastError());
}

static int fe_error(afe_space *s, const double *d_uh, int acc, int h1, double *err) {
  API_BEGIN
  if (!s || !d_uh || !err) return fail(AFE_E_ARG, "null argument");
  if (!s->ctx) return fail(AFE_E_ARG, "host-only space (no device context)");
  const QuadRule *q = quad_find(s->h.tdim, acc);
  if (!q) return fail(AFE_E_ARG, "no quadrature rule of that accuracy");
  std::lock_guard<CtxMutex> lk(s->ctx->mu);
  HIPCHK(hipSetDevice(s->ctx->device));
  DevBuf<double> pe, part, tot;
  pe.alloc(s->ctx, s->h.n_elem);
  part.alloc(s->ctx, RED_BLOCKS);
  tot.alloc(s->ctx, 1);
  switch (s->h.kind) {
  case P1_TRI: launch_error<P1_TRI>(s, q, h1, d_uh, pe.p); break;
  case P2_TRI: launch_error<P2_TRI>(s, q, h1, d_uh, pe.p); break;
  default: launch_error<P1_TET>(s, q, h1, d_uh, pe.p); break;
  }
  k_sum_partial<<<RED_BLOCKS, 256, 0, s->ctx->stream>>>(s->h.n_elem, pe.p, part.p);
  k_finalize<<<1, FIN_THREADS, 0, s->ctx->stream>>>(part.p, RED_BLOCKS, 1, tot.p);
  double h = 0;
  HIPCHK(hipMemcpyAsync(&h, tot.p, sizeof(double), hipMemcpyDeviceToHost, s->ctx->stream));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  *err = std::sqrt(std::fabs(h));
  API_END
}

extern "C" int afe_error_l2(afe_space *s, const double *d_uh, int32_t acc, double *err) {
  return fe_error(s, d_uh, acc, 0, err);
}
extern "C" int afe_error_h1(afe_space *s, const double *d_uh, int32_t acc, double *err) {
  return fe_error(s, d_uh, acc, 1, err);
}

// ===========================================================================
// AMG hierarchy on the device
// ===========================================================================
constexpr int64_t COARSE_TAIL_MAX = 256;

struct DevLevel {
  int64_t n = 0;
  int ncol = 0;
  std::vector<int32_t> colour_slice; // host copy
  int64_t rows_c0 = 0;               // rows of colour 0 (first in the colour order)
  std::vector<int32_t> colour_row;   // host copy: rows of colour c are [colour_row[c], colour_row[c+1])
  int ns = 0, p_ns = 0, t_ns = 0;
  int gs = 0, p_gs = 0; // lanes per row (shift) of the level matrix and of the restriction
  int zs = 0;           // first slice of the last colour
  DevBuf<SliceInfo> si, p_si, t_si;
  DevBuf<int32_t> cols, p_cols, t_cols, col_slice;
  DevBuf<uint16_t> c16; // colour-coded 16-bit columns (streamed level 0), used by k_gs / k_residual
  DevBuf<int32_t> cb16;
  int cshift = 0;
  DevBuf<double> v64, invd64, diag64, p_v64, t_v64;
  DevBuf<float> v32, invd32, diag32, p_v32, t_v32;
  DevBuf<double> x64, b64, t64;
  DevBuf<float> x32, b32, t32;
  std::vector<int32_t> perm; // new -> old
  bool stream = false;       // matrix larger than the caches: non-temporal loads
  // value-only refresh of a lazy hierarchy: the level's CSR (levels >= 1),
  // the restriction P and its transpose to the next level, and the source
  // maps CSR entry -> layout position / stored row -> CSR diagonal
  DevBuf<int64_t> c_rp, P_rp, Pt_rp;
  DevBuf<int32_t> c_col, P_col, Pt_col, smap, dsrc;
  DevBuf<double> c_val, P_val, Pt_val;
};

struct afe_amg {
  afe_ctx *ctx = nullptr;
  afe_matrix *A0 = nullptr;
  // the pattern the hierarchy was built on, for the same-pattern rule of
  // lazyReinit: the space of a space matrix, else the CSR matrix object.
  // Kept apart from A0 so that the check never reads a matrix the caller
  // has destroyed since.
  const afe_space *key_space = nullptr;
  const afe_matrix *key_mat = nullptr;
  uint64_t graph_epoch = 0; // g_tune.epoch when the graphs below were captured
  // live timing of the CG's level-0 SpMV inside the solves (events recorded
  // around it in every iteration, summed after each iteration's sync)
  bool live = false;
  hipEvent_t live0 = nullptr, live1 = nullptr;
  double live_ms = 0;
  int64_t live_n = 0, live_missed = 0;
  // part timing (afe_amg_part_timing): events before part 0, between the
  // parts and after part 1 of a CG iteration, in two alternating sets (an
  // iteration's set is read once the next iteration's part 0 has finished)
  bool parts = false, part_pending = false;
  int part_set = 0;
  hipEvent_t pev[2][3] = {};
  double part_ms[2] = {0, 0};
  int64_t part_n = 0;
  afe_amg_params prm{};
  Hierarchy h;
  std::vector<DevLevel> lv;
  int tail0 = 0;       // first level handled by the single-workgroup tail
  int coarse_mode = 0; // 0 dense inverse, 1 diagonal, 2 GS
  bool coarse_exact = true; // AMGSolver::isSolveMostProjectExactly()
  bool big_coarse = false; // dense coarsest level solved by k_coarse_gemv
  DevBuf<double> minv;
  DevBuf<LevelDev<double>> ld64;
  DevBuf<LevelDev<float>> ld32;
  std::vector<LevelDev<double>> h_ld64; // host copies of the level descriptors
  std::vector<LevelDev<float>> h_ld32;
  DevBuf<int32_t> perm0, iperm0; // colour index -> natural row, and back
  // level-0 FP64 matrix in the natural numbering (the CG SpMV and the true
  // residual): rows sorted by stored length inside windows of NAT_WINDOW
  // rows, nat_rowmap = stored row -> natural row
  int nat_ns = 0;
  DevBuf<SliceInfo> nat_si;
  DevBuf<int32_t> nat_cols, nat_rowmap, nat_smap, nat_dsrc;
  DevBuf<uint16_t> nat_c16; // 16-bit columns (with nat_cbase per slice) in place of nat_cols
  DevBuf<int32_t> nat_cbase;
  DevBuf<double> nat_v64, nat_diag;
  bool dev_refresh = false;      // device lazyReinit available (source maps uploaded)
  bool host_vals_stale = false;  // host level values older than the device's
  // CG workspace (level-0 natural numbering)
  DevBuf<double> wx, wxl, wr, wp, wq, wz, wb, part;
  DevBuf<CgScalars> sc;
  CgScalars *h_sc = nullptr; // pinned
  bool has32 = false;
  // captured V-cycles: [precision][symmetric]
  hipGraphExec_t vc_graph[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  // captured CG iteration in two parts: [part][precision]; part 0 ends with
  // the copy of the scalars to the host, part 1 is the preconditioner and the
  // direction update
  hipGraphExec_t cg_graph[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  int graph_nu[2][2] = {{-1, -1}, {-1, -1}};
  int cg_nu[2][2] = {{-1, -1}, {-1, -1}};
  hipEvent_t it_ev = nullptr; // recorded after part 0 of each CG iteration
};

static void drop_graphs(afe_amg *g) {
  bool any = false;
  for (auto &row : g->vc_graph)
    for (auto &e : row) any = any || e;
  for (auto &row : g->cg_graph)
    for (auto &e : row) any = any || e;
  // no executable graph is destroyed while a launch of it may still run
  if (any && g->ctx) (void)hipStreamSynchronize(g->ctx->stream);
  for (auto &row : g->vc_graph)
    for (auto &e : row)
      if (e) {
        (void)hipGraphExecDestroy(e);
        e = nullptr;
      }
  for (auto &row : g->cg_graph)
    for (auto &e : row)
      if (e) {
        (void)hipGraphExecDestroy(e);
        e = nullptr;
      }
}

// host CSR of the fine matrix (values from the device)
static Csr fine_csr(afe_matrix *a) {
  Csr c;
  c.m = a->m;
  c.n = a->n;
  c.rp = a->hrp();
  c.col = a->hcol();
  c.val.resize(a->nnz);
  HIPCHK(hipMemcpyAsync(c.val.data(), a->val.p, a->nnz * sizeof(double), hipMemcpyDeviceToHost, a->ctx->stream));
  HIPCHK(hipStreamSynchronize(a->ctx->stream));
  return c;
}

template <class T> static std::vector<T> to_prec(const std::vector<double> &v) {
  return std::vector<T>(v.begin(), v.end());
}

static std::vector<double> inv_of(const std::vector<double> &d) {
  std::vector<double> r(d.size());
  for (size_t i = 0; i < d.size(); i++) r[i] = 1.0 / d[i];
  return r;
}

// colour, permute and convert every level of the host hierarchy; upload
// device layout of a hierarchy, computed on the host without touching the
// device (so that it can be prepared while other work uses the context)
struct HostLevel {
  int ncol = 0;
  int64_t n = 0;
  Sell sa, sp, st;
  // colour-coded 16-bit columns of a streamed level (empty when they do not fit)
  std::vector<uint16_t> c16;
  std::vector<int32_t> cb16;
  int cshift = 0;
  std::vector<int32_t> perm, iperm;
  std::vector<int32_t> smap, dsrc; // CSR entry -> layout position, row -> CSR diagonal
};
struct HostLayout {
  bool maps = false; // source maps for the value-only refresh
  std::vector<HostLevel> lv;
  Sell nat;
  std::vector<int32_t> nat_rowmap, nat_smap, nat_dsrc;
  // 16-bit columns of the natural copy: offsets from a base column per
  // slice, when every slice's columns span at most 65535 (else empty)
  std::vector<uint16_t> nat_c16;
  std::vector<int32_t> nat_cbase;
  std::vector<double> minv_perm;
};

// colour-coded 16-bit columns of a level's layout (see sell_dot_cc): per
// slice and colour the smallest position of that colour among the slice's
// columns is the base; kept only when every offset fits in 16 - colour bits
static void colour_code(HostLevel &hl) {
  const Sell &sa = hl.sa;
  const int cbits = hl.ncol <= 2 ? 1 : (hl.ncol <= 4 ? 2 : 3);
  const int cshift = 16 - cbits;
  const int32_t maxoff = (1 << cshift) - 1;
  std::vector<uint8_t> pc(sa.n_rows);
  for (int c = 0; c < hl.ncol; c++)
    for (int32_t p = sa.group_row[c]; p < sa.group_row[c + 1]; p++) pc[p] = (uint8_t)c;
  const int64_t ns = (int64_t)sa.slices.size();
  std::vector<int32_t> cb((size_t)ns * CC_BASES, 0);
  std::vector<uint16_t> c16(sa.cols.size(), 0);
  for (int64_t q = 0; q < ns; q++) {
    const SliceInfo &si = sa.slices[q];
    int32_t lo[CC_BASES], hi[CC_BASES];
    for (int c = 0; c < CC_BASES; c++) {
      lo[c] = INT32_MAX;
      hi[c] = INT32_MIN;
    }
    const int64_t lanes = (int64_t)si.nrow << sa.gshift;
    for (int64_t j = 0; j < si.width; j++)
      for (int64_t l = 0; l < lanes; l++) {
        const int32_t col = sa.cols[si.off + j * SELL_C + l];
        const int c = pc[col];
        lo[c] = std::min(lo[c], col);
        hi[c] = std::max(hi[c], col);
      }
    for (int c = 0; c < CC_BASES; c++) {
      if (hi[c] < lo[c]) continue;
      if (hi[c] - lo[c] > maxoff) return; // does not fit: keep the int32 columns only
      cb[(size_t)q * CC_BASES + c] = lo[c];
    }
    for (int64_t j = 0; j < si.width; j++)
      for (int64_t l = 0; l < lanes; l++) {
        const int64_t e = si.off + j * SELL_C + l;
        const int c = pc[sa.cols[e]];
        c16[e] = (uint16_t)(c << cshift | (sa.cols[e] - cb[(size_t)q * CC_BASES + c]));
      }
  }
  hl.c16 = std::move(c16);
  hl.cb16 = std::move(cb);
  hl.cshift = cshift;
}

static void build_layout(const Hierarchy &h, HostLayout &lay, bool maps) {
  const int L = (int)h.A.size();
  lay.maps = maps && h.A[0].nnz() < INT32_MAX;
  lay.lv.assign(L, HostLevel());
  std::vector<std::vector<int32_t>> groups(L);
  // phase times for AFE_VERBOSE: colouring, row order, level sliced ELL,
  // transfer operators, natural copy
  double tph[5] = {0, 0, 0, 0, 0};
  auto tick = [t = std::chrono::steady_clock::now()](double &acc) mutable {
    const auto now = std::chrono::steady_clock::now();
    acc += std::chrono::duration<double>(now - t).count();
    t = now;
  };
  double tdrop = 0;
  tick(tdrop);
  for (int l = 0; l < L; l++) {
    const Csr &A = h.A[l];
    HostLevel &hl = lay.lv[l];
    std::vector<int32_t> colour;
    const int ncol = colour_rows(A, colour, t_first_fit_only ? 0 : l == 0 ? DSATUR_MAX_ROWS : DSATUR_MAX_ROWS_COARSE);
    tick(tph[0]);
    std::vector<int32_t> len(A.m, 0);
    for (int64_t i = 0; i < A.m; i++)
      for (int64_t k = A.rp[i]; k < A.rp[i + 1]; k++)
        if (A.col[k] != i) len[i]++;
    // rows of one colour keep their index order in windows of COLOUR_WINDOW
    // rows (x locality of the gathers) and are sorted by stored length inside
    // a window (little sliced-ELL padding); ties keep the index order.  A
    // counting sort by colour, then a stable sort of each window.
    std::vector<int64_t> cstart(ncol + 1, 0);
    for (int64_t i = 0; i < A.m; i++) cstart[colour[i] + 1]++;
    for (int c = 0; c < ncol; c++) cstart[c + 1] += cstart[c];
    hl.perm.resize(A.m);
    {
      std::vector<int64_t> fill(cstart.begin(), cstart.end() - 1);
      for (int64_t i = 0; i < A.m; i++) hl.perm[fill[colour[i]]++] = (int32_t)i;
    }
    for (int c = 0; c < ncol; c++)
      for (int64_t w = cstart[c]; w < cstart[c + 1]; w += COLOUR_WINDOW)
        std::stable_sort(hl.perm.begin() + w, hl.perm.begin() + std::min<int64_t>(cstart[c + 1], w + COLOUR_WINDOW),
                         [&](int32_t a, int32_t b) { return len[a] > len[b]; });
    hl.iperm.resize(A.m);
    for (int64_t i = 0; i < A.m; i++) hl.iperm[hl.perm[i]] = (int32_t)i;
    groups[l].assign(ncol, 0);
    for (int64_t i = 0; i < A.m; i++) groups[l][colour[i]]++;
    hl.ncol = ncol;
    hl.n = A.m;
    tick(tph[1]);
  }
  // lanes per row: level 0 keeps one lane per row (its colours fill the
  // chip), the other levels and their restrictions follow row_lane_shift.
  // AFE_ROW_LANES=G (1, 2, 4, 8 or 16) sets G on every level >= 1 instead
  // (G = 1: one lane per row everywhere; the tests compare the forms)
  int forced = -1;
  if (const char *e = getenv("AFE_ROW_LANES")) {
    const int v = atoi(e);
    forced = 0;
    while (forced < MAX_ROW_LANE_SHIFT && (2 << forced) <= v) forced++;
  }
  auto lane_shift = [&](int64_t rows, int64_t groups, int64_t entries) {
    return forced >= 0 ? forced : row_lane_shift(rows, groups, entries, MAX_ROW_LANE_SHIFT);
  };
  for (int l = 0; l < L; l++) {
    HostLevel &hl = lay.lv[l];
    const Csr &A = h.A[l];
    const int gs = l == 0 ? 0 : lane_shift(A.m, hl.ncol, A.nnz() - A.m);
    build_sell(A, hl.perm, hl.iperm, groups[l], true, hl.sa, lay.maps ? &hl.smap : nullptr,
               lay.maps ? &hl.dsrc : nullptr, gs);
    if ((int64_t)hl.sa.cols.size() * 8 > stream_bytes() && hl.ncol <= CC_BASES && getenv("AFE_NO_C16") == nullptr)
      colour_code(hl);
    tick(tph[2]);
    if (l + 1 < L) {
      build_sell(h.P[l], lay.lv[l + 1].perm, hl.iperm, {(int32_t)h.P[l].m}, false, hl.sp, nullptr, nullptr,
                 l == 0 ? 0 : lane_shift(h.P[l].m, 1, h.P[l].nnz()));
      // prolongation is P^T of the restriction (SparseMatrix::Tvmult of P in
      // AMGSolver::solve, src/AMGSolver.cpp:608-613), stored explicitly
      build_sell(transpose(h.P[l]), hl.perm, lay.lv[l + 1].iperm, {(int32_t)h.A[l].m}, false, hl.st);
      tick(tph[3]);
    }
  }
  {
    // level 0 in the natural numbering for the CG
    const Csr &A = h.A[0];
    std::vector<int32_t> len(A.m, 0), ident(A.m);
    for (int64_t i = 0; i < A.m; i++)
      for (int64_t k = A.rp[i]; k < A.rp[i + 1]; k++)
        if (A.col[k] != i) len[i]++;
    lay.nat_rowmap.resize(A.m);
    std::iota(lay.nat_rowmap.begin(), lay.nat_rowmap.end(), 0);
    std::iota(ident.begin(), ident.end(), 0);
    for (int64_t w = 0; w < A.m; w += NAT_WINDOW)
      std::stable_sort(lay.nat_rowmap.begin() + w, lay.nat_rowmap.begin() + std::min<int64_t>(A.m, w + NAT_WINDOW),
                       [&](int32_t a, int32_t b) { return len[a] > len[b]; });
    build_sell(A, lay.nat_rowmap, ident, {(int32_t)A.m}, true, lay.nat, lay.maps ? &lay.nat_smap : nullptr,
               lay.maps ? &lay.nat_dsrc : nullptr);
    // rows of a slice lie in one window of NAT_WINDOW rows, so their columns
    // span little more than the matrix bandwidth: 16-bit offsets from the
    // slice's smallest column when that fits
    // (AFE_NO_C16=1 keeps the 32-bit columns: both forms are tested against
    // each other)
    const int64_t ns = (int64_t)lay.nat.slices.size();
    std::vector<int32_t> base(ns);
    bool fits = getenv("AFE_NO_C16") == nullptr;
    // (the lanes of a partial last slice hold no row: they stay 0 and are
    // never read)
    for (int64_t q = 0; q < ns && fits; q++) {
      const SliceInfo &si = lay.nat.slices[q];
      int32_t lo = INT32_MAX, hi = INT32_MIN;
      for (int64_t j = 0; j < si.width; j++)
        for (int64_t l = 0; l < si.nrow; l++) {
          const int32_t c = lay.nat.cols[si.off + j * SELL_C + l];
          lo = std::min(lo, c);
          hi = std::max(hi, c);
        }
      if (si.width == 0 || si.nrow == 0) lo = hi = 0;
      base[q] = lo;
      fits = (int64_t)hi - lo <= 65535;
    }
    if (fits) {
      lay.nat_c16.assign(lay.nat.cols.size(), 0);
      for (int64_t q = 0; q < ns; q++) {
        const SliceInfo &si = lay.nat.slices[q];
        for (int64_t j = 0; j < si.width; j++)
          for (int64_t l = 0; l < si.nrow; l++) {
            const int64_t e = si.off + j * SELL_C + l;
            lay.nat_c16[e] = (uint16_t)(lay.nat.cols[e] - base[q]);
          }
      }
      lay.nat_cbase = std::move(base);
    }
    tick(tph[4]);
  }
  if (verbose())
    fprintf(stderr, "layout: colouring %.3f s, row order %.3f s, level ELL %.3f s, transfer %.3f s, natural %.3f s\n",
            tph[0], tph[1], tph[2], tph[3], tph[4]);
  if (h.full) {
    // the coarse vectors are in the level's colour order: permute the inverse
    const std::vector<int32_t> &pc = lay.lv[L - 1].perm;
    const int64_t nc = (int64_t)pc.size();
    lay.minv_perm.resize((size_t)nc * nc);
    for (int64_t i = 0; i < nc; i++)
      for (int64_t j = 0; j < nc; j++) lay.minv_perm[(size_t)i * nc + j] = h.minv[(size_t)pc[i] * nc + pc[j]];
  }
}

// uploads a prepared layout (caller holds the context lock)
// coarsest-level solve (src/AMGSolver.cpp:590-605): the dense inverse when the
// coarsening ended on a full matrix, diagonal scaling when it stalled, and
// Gauss-Seidel sweeps from zero (multicolour, forward colour order, smoothing
// count per side) when the caller turned the exact solve off
static void set_coarse_mode(afe_amg *g) {
  const int L = (int)g->lv.size();
  if (!g->coarse_exact) {
    g->coarse_mode = 2;
    g->big_coarse = false;
  } else {
    g->coarse_mode = g->h.full ? 0 : 1;
    g->big_coarse = g->h.full && g->lv[L - 1].n > COARSE_TAIL_MAX;
  }
}

static void finish_levels(afe_amg *g);

static void upload_layout(afe_amg *g, HostLayout &lay) {
  afe_ctx *ctx = g->ctx;
  const Hierarchy &h = g->h;
  const int L = (int)h.A.size();
  drop_graphs(g);
  g->lv.clear();
  g->lv.resize(L);
  for (int l = 0; l < L; l++) {
    DevLevel &d = g->lv[l];
    HostLevel &hl = lay.lv[l];
    d.ncol = hl.ncol;
    d.n = hl.n;
    Sell &sa = hl.sa;
    d.ns = (int)sa.slices.size();
    d.gs = sa.gshift;
    d.colour_slice = sa.group_slice;
    d.zs = d.ncol > 0 ? sa.group_slice[d.ncol - 1] : d.ns;
    d.rows_c0 = sa.group_row.size() > 1 ? sa.group_row[1] : hl.n;
    d.colour_row = sa.group_row;
    d.si.upload(ctx, sa.slices);
    d.cols.upload(ctx, sa.cols);
    d.col_slice.upload(ctx, sa.group_slice);
    if (!hl.c16.empty()) {
      d.c16.upload(ctx, hl.c16);
      d.cb16.upload(ctx, hl.cb16);
      d.cshift = hl.cshift;
    }
    d.v64.upload(ctx, sa.vals);
    d.diag64.upload(ctx, sa.diag);
    d.invd64.upload(ctx, inv_of(sa.diag));
    if (g->has32) {
      d.v32.upload(ctx, to_prec<float>(sa.vals));
      d.diag32.upload(ctx, to_prec<float>(sa.diag));
      d.invd32.upload(ctx, to_prec<float>(inv_of(sa.diag)));
    }
    if (l + 1 < L) {
      Sell &sp = hl.sp, &st = hl.st;
      d.p_ns = (int)sp.slices.size();
      d.p_gs = sp.gshift;
      d.t_ns = (int)st.slices.size();
      d.p_si.upload(ctx, sp.slices);
      d.p_cols.upload(ctx, sp.cols);
      d.p_v64.upload(ctx, sp.vals);
      d.t_si.upload(ctx, st.slices);
      d.t_cols.upload(ctx, st.cols);
      d.t_v64.upload(ctx, st.vals);
      if (g->has32) {
        d.p_v32.upload(ctx, to_prec<float>(sp.vals));
        d.t_v32.upload(ctx, to_prec<float>(st.vals));
      }
    }
    d.x64.alloc(ctx, d.n);
    d.b64.alloc(ctx, d.n);
    d.t64.alloc(ctx, d.n);
    if (g->has32) {
      d.x32.alloc(ctx, d.n);
      d.b32.alloc(ctx, d.n);
      d.t32.alloc(ctx, d.n);
    }
    d.perm = hl.perm;
    // stream (non-temporal) when the level's FP32 matrix exceeds the cache
    d.stream = (int64_t)sa.cols.size() * 8 > stream_bytes();
  }
  g->perm0.upload(ctx, lay.lv[0].perm);
  g->iperm0.upload(ctx, lay.lv[0].iperm);
  g->nat_ns = (int)lay.nat.slices.size();
  g->nat_si.upload(ctx, lay.nat.slices);
  if (!lay.nat_c16.empty()) {
    g->nat_c16.upload(ctx, lay.nat_c16);
    g->nat_cbase.upload(ctx, lay.nat_cbase);
    g->nat_cols.release();
  } else {
    g->nat_cols.upload(ctx, lay.nat.cols);
    g->nat_c16.release();
    g->nat_cbase.release();
  }
  g->nat_v64.upload(ctx, lay.nat.vals);
  g->nat_diag.upload(ctx, lay.nat.diag);
  g->nat_rowmap.upload(ctx, lay.nat_rowmap);
  g->dev_refresh = lay.maps;
  g->host_vals_stale = false;
  if (lay.maps) {
    g->nat_smap.upload(ctx, lay.nat_smap);
    g->nat_dsrc.upload(ctx, lay.nat_dsrc);
    for (int l = 0; l < L; l++) {
      DevLevel &d = g->lv[l];
      d.smap.upload(ctx, lay.lv[l].smap);
      d.dsrc.upload(ctx, lay.lv[l].dsrc);
      if (l > 0) {
        d.c_rp.upload(ctx, h.A[l].rp);
        d.c_col.upload(ctx, h.A[l].col);
        d.c_val.upload(ctx, h.A[l].val);
      }
      if (l + 1 < L) {
        d.P_rp.upload(ctx, h.P[l].rp);
        d.P_col.upload(ctx, h.P[l].col);
        d.P_val.upload(ctx, h.P[l].val);
        d.Pt_rp.upload(ctx, h.Pt[l].rp);
        d.Pt_col.upload(ctx, h.Pt[l].col);
        d.Pt_val.upload(ctx, h.Pt[l].val);
      }
    }
  }
  // first level of the coarse tail; a large dense coarsest level is solved by
  // the multi-block gemv between the tail's down and up legs
  set_coarse_mode(g);
  g->tail0 = L - 1;
  while (g->tail0 > 0 && g->lv[g->tail0 - 1].n <= g_tune.tail_rows) g->tail0--;
  if (h.full) g->minv.upload(ctx, lay.minv_perm);
  finish_levels(g);
}

// the level descriptors of the tail and batch kernels (device pointers into
// the levels' buffers) and the CG work vectors
static void finish_levels(afe_amg *g) {
  afe_ctx *ctx = g->ctx;
  const int L = (int)g->lv.size();
  std::vector<LevelDev<double>> d64(L);
  std::vector<LevelDev<float>> d32(L);
  for (int l = 0; l < L; l++) {
    DevLevel &d = g->lv[l];
    LevelDev<double> a{};
    a.n = (int)d.n; a.ncol = d.ncol; a.ns = d.ns; a.p_ns = d.p_ns; a.t_ns = d.t_ns;
    a.gs = d.gs; a.p_gs = d.p_gs; a.zs = d.zs; a.rows_c0 = (int)d.rows_c0;
    a.rows_lc = d.ncol > 0 ? d.colour_row[d.ncol - 1] : (int)d.n;
    a.nst = (int)d.cols.n;
    a.col_slice = d.col_slice.p; a.si = d.si.p; a.cols = d.cols.p;
    a.vals = d.v64.p; a.invd = d.invd64.p; a.diag = d.diag64.p;
    a.p_si = d.p_si.p; a.p_cols = d.p_cols.p; a.p_vals = d.p_v64.p;
    a.t_si = d.t_si.p; a.t_cols = d.t_cols.p; a.t_vals = d.t_v64.p;
    a.x = d.x64.p; a.b = d.b64.p; a.t = d.t64.p;
    d64[l] = a;
    LevelDev<float> b{};
    b.n = a.n; b.ncol = a.ncol; b.ns = a.ns; b.p_ns = a.p_ns; b.t_ns = a.t_ns;
    b.gs = a.gs; b.p_gs = a.p_gs; b.zs = a.zs; b.rows_c0 = a.rows_c0; b.rows_lc = a.rows_lc; b.nst = a.nst;
    b.col_slice = a.col_slice; b.si = a.si; b.cols = a.cols;
    b.vals = d.v32.p; b.invd = d.invd32.p; b.diag = d.diag32.p;
    b.p_si = a.p_si; b.p_cols = a.p_cols; b.p_vals = d.p_v32.p;
    b.t_si = a.t_si; b.t_cols = a.t_cols; b.t_vals = d.t_v32.p;
    b.x = d.x32.p; b.b = d.b32.p; b.t = d.t32.p;
    d32[l] = b;
  }
  g->ld64.upload(ctx, d64);
  if (g->has32) g->ld32.upload(ctx, d32);
  g->h_ld64 = d64;
  g->h_ld32 = d32;
  const int64_t n0 = g->lv[0].n;
  g->wx.alloc(ctx, n0);
  g->wxl.alloc(ctx, n0);
  g->wr.alloc(ctx, n0);
  g->wp.alloc(ctx, n0);
  g->wq.alloc(ctx, n0);
  g->wz.alloc(ctx, n0);
  g->wb.alloc(ctx, n0);
  g->part.alloc(ctx, 3 * (size_t)kMaxGridBlocks);
  g->sc.alloc(ctx, 1);
  HIPCHK(hipStreamSynchronize(ctx->stream));
}

extern "C" int afe_amg_create(afe_matrix *a, const afe_amg_params *p, afe_amg **out) {
  API_BEGIN
  if (!a || !out) return fail(AFE_E_ARG, "null argument");
  if (a->m != a->n || a->ps0) return fail(AFE_E_ARG, "AMG needs the square matrix of one space");
  auto g = std::make_unique<afe_amg>();
  g->ctx = a->ctx;
  g->A0 = a;
  g->key_space = a->s;
  g->key_mat = a->s ? nullptr : a;
  afe_amg_params prm{0.25, 50u, 0.382, 0, 3, 1};
  if (p) prm = *p;
  if (prm.smooth < 0 || prm.nmpo < 1) return fail(AFE_E_ARG, "bad AMG parameters");
  g->prm = prm;
  g->has32 = prm.build_fp32 != 0;
  AmgParams hp;
  hp.alpha = prm.alpha;
  hp.nmpo = prm.nmpo;
  hp.nmpsd = prm.nmpsd;
  hp.lazy = prm.lazy;
  Csr A0;
  {
    std::lock_guard<CtxMutex> lk(a->ctx->mu);
    HIPCHK(hipSetDevice(a->ctx->device));
    A0 = fine_csr(a);
  }
  // the coarsening and the device layout are host work: the context stays
  // free for other callers meanwhile
  const auto t0 = std::chrono::steady_clock::now();
  amg_build(A0, hp, g->h);
  const auto t1 = std::chrono::steady_clock::now();
  HostLayout lay;
  build_layout(g->h, lay, prm.lazy != 0);
  const auto t2 = std::chrono::steady_clock::now();
  std::lock_guard<CtxMutex> lk(a->ctx->mu);
  HIPCHK(hipSetDevice(a->ctx->device));
  HIPCHK(hipHostMalloc((void **)&g->h_sc, sizeof(CgScalars), hipHostMallocDefault));
  upload_layout(g.get(), lay);
  if (verbose()) {
    const auto t3 = std::chrono::steady_clock::now();
    fprintf(stderr, "amg_create n %lld: coarsening %.3f s, layout %.3f s, upload %.3f s\n", (long long)A0.m,
            std::chrono::duration<double>(t1 - t0).count(), std::chrono::duration<double>(t2 - t1).count(),
            std::chrono::duration<double>(t3 - t2).count());
  }
  *out = g.release();
  API_END
}

extern "C" int afe_amg_create_host(int64_t m, const int64_t *row_ptr, const int64_t *col, const double *val,
                                   const afe_amg_params *p, afe_amg **out) {
  API_BEGIN
  if (!row_ptr || !col || !val || !out || m <= 0) return fail(AFE_E_ARG, "null argument");
  auto g = std::make_unique<afe_amg>();
  afe_amg_params prm{0.25, 50u, 0.382, 0, 3, 0};
  if (p) prm = *p;
  g->prm = prm;
  AmgParams hp;
  hp.alpha = prm.alpha;
  hp.nmpo = prm.nmpo;
  hp.nmpsd = prm.nmpsd;
  hp.lazy = prm.lazy;
  Csr A0;
  A0.m = A0.n = m;
  A0.rp.assign(row_ptr, row_ptr + m + 1);
  A0.col.resize(row_ptr[m]);
  for (int64_t k = 0; k < row_ptr[m]; k++) {
    if (col[k] < 0 || col[k] >= m) return fail(AFE_E_ARG, "column index out of range");
    A0.col[k] = (int32_t)col[k];
  }
  A0.val.assign(val, val + row_ptr[m]);
  amg_build(A0, hp, g->h);
  *out = g.release();
  API_END
}

extern "C" int afe_amg_lazy_reinit_host(afe_amg *g, const double *val) {
  API_BEGIN
  if (!g || !val || g->ctx) return fail(AFE_E_ARG, "not a host-only hierarchy");
  if (!g->prm.lazy) return fail(AFE_E_ARG, "hierarchy was not built in lazy mode");
  Csr A0 = g->h.A[0];
  std::memcpy(A0.val.data(), val, sizeof(double) * A0.nnz());
  amg_lazy_refresh(g->h, A0);
  API_END
}

extern "C" int afe_amg_coarse_inverse(const afe_amg *g, double *minv) {
  if (!g || !minv) return fail(AFE_E_ARG, "null argument");
  if (!g->h.full) return fail(AFE_E_ARG, "coarsest level is not solved by a dense inverse");
  std::memcpy(minv, g->h.minv.data(), sizeof(double) * g->h.minv.size());
  return AFE_OK;
}

extern "C" int afe_amg_destroy(afe_amg *g) {
  if (!g) return AFE_OK;
  if (!g->ctx) {
    delete g;
    return AFE_OK;
  }
  std::lock_guard<CtxMutex> lk(g->ctx->mu);
  (void)hipSetDevice(g->ctx->device);
  (void)hipStreamSynchronize(g->ctx->stream);
  drop_graphs(g);
  if (g->h_sc) (void)hipHostFree(g->h_sc);
  if (g->live0) (void)hipEventDestroy(g->live0);
  if (g->live1) (void)hipEventDestroy(g->live1);
  for (auto &set : g->pev)
    for (auto &e : set)
      if (e) (void)hipEventDestroy(e);
  if (g->it_ev) (void)hipEventDestroy(g->it_ev);
  delete g;
  return AFE_OK;
}

// AMGSolver::lazyReinit on the device: Galerkin products level by level with
// the kept P (bit-identical to the host product), the layouts' values through
// their source maps, and the coarsest inverse on the host from the new
// coarsest values.  Layouts, launch tables and captured graphs stay valid.
static void device_lazy_refresh(afe_amg *g, afe_matrix *a) {
  afe_ctx *ctx = g->ctx;
  std::lock_guard<CtxMutex> lk(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  Hierarchy &h = g->h;
  const int L = (int)g->lv.size();
  auto lvals = [&](int l) -> const double * { return l == 0 ? a->val.p : g->lv[l].c_val.p; };
  for (int l = 0; l + 1 < L; l++) {
    DevLevel &d = g->lv[l];
    DevLevel &dn = g->lv[l + 1];
    const int64_t *arp = l == 0 ? a->rp() : d.c_rp.p;
    const int32_t *acol = l == 0 ? a->col() : d.c_col.p;
    const int64_t nc = dn.n;
    k_galerkin_rows<<<nblk(nc), 256, 0, s>>>(nc, d.P_rp.p, d.P_col.p, d.P_val.p, arp, acol, lvals(l), d.Pt_rp.p,
                                              d.Pt_col.p, d.Pt_val.p, dn.c_rp.p, dn.c_col.p, dn.c_val.p);
  }
  for (int l = 0; l < L; l++) {
    DevLevel &d = g->lv[l];
    const int64_t nnz = h.A[l].nnz();
    k_refresh_vals<<<nblk(nnz), 256, 0, s>>>(nnz, lvals(l), d.smap.p, d.v64.p, g->has32 ? d.v32.p : nullptr);
    k_refresh_diag<<<nblk(d.n), 256, 0, s>>>(d.n, lvals(l), d.dsrc.p, d.diag64.p, d.invd64.p,
                                             g->has32 ? d.diag32.p : nullptr, g->has32 ? d.invd32.p : nullptr);
  }
  const int64_t n0 = g->lv[0].n, nnz0 = h.A[0].nnz();
  k_refresh_vals<<<nblk(nnz0), 256, 0, s>>>(nnz0, a->val.p, g->nat_smap.p, g->nat_v64.p, nullptr);
  k_refresh_diag<<<nblk(n0), 256, 0, s>>>(n0, a->val.p, g->nat_dsrc.p, g->nat_diag.p, nullptr, nullptr, nullptr);
  HIPCHK(hipGetLastError());
  if (h.full) {
    Csr &C = h.A[L - 1];
    HIPCHK(hipMemcpyAsync(C.val.data(), lvals(L - 1), C.nnz() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    coarse_inverse(h);
    const std::vector<int32_t> &pc = g->lv[L - 1].perm;
    const int64_t nc = (int64_t)pc.size();
    std::vector<double> mp((size_t)nc * nc);
    for (int64_t i = 0; i < nc; i++)
      for (int64_t j = 0; j < nc; j++) mp[(size_t)i * nc + j] = h.minv[(size_t)pc[i] * nc + pc[j]];
    g->minv.upload(ctx, mp);
  }
  HIPCHK(hipStreamSynchronize(s));
  g->A0 = a;
  g->host_vals_stale = true;
}

// host copies of the level values after device refreshes
static void sync_host_values(afe_amg *g) {
  if (!g->host_vals_stale || !g->ctx) return;
  std::lock_guard<CtxMutex> lk(g->ctx->mu);
  hipStream_t s = g->ctx->stream;
  for (size_t l = 0; l < g->lv.size(); l++) {
    Csr &A = g->h.A[l];
    const double *src = l == 0 ? g->A0->val.p : g->lv[l].c_val.p;
    HIPCHK(hipMemcpyAsync(A.val.data(), src, A.nnz() * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  HIPCHK(hipStreamSynchronize(s));
  g->host_vals_stale = false;
}

extern "C" int afe_amg_lazy_reinit(afe_amg *g, afe_matrix *a) {
  API_BEGIN
  if (!g || !a) return fail(AFE_E_ARG, "null argument");
  if (!g->prm.lazy) return fail(AFE_E_ARG, "hierarchy was not built in lazy mode");
  if (a->s != g->key_space || (a->s == nullptr && a != g->key_mat))
    return fail(AFE_E_PATTERN, "The solver is not initialized with the same sparsity pattern.");
  if (g->dev_refresh) {
    device_lazy_refresh(g, a);
    return AFE_OK;
  }
  Csr A0;
  {
    std::lock_guard<CtxMutex> lk(g->ctx->mu);
    HIPCHK(hipSetDevice(g->ctx->device));
    A0 = fine_csr(a);
  }
  // host Galerkin products with the kept P and the new layout, unlocked (the
  // caller must not use this hierarchy concurrently)
  amg_lazy_refresh(g->h, A0);
  HostLayout lay;
  build_layout(g->h, lay, true);
  std::lock_guard<CtxMutex> lk(g->ctx->mu);
  HIPCHK(hipSetDevice(g->ctx->device));
  g->A0 = a;
  upload_layout(g, lay);
  API_END
}

// ---------------------------------------------------------------------------
// structure sharing: the numbering, pattern and gather map of a space, and the
// interpolation, colouring and layouts of a lazy hierarchy, depend on the
// mesh connectivity only (lazyProject uses the pattern, not the values), so a
// mesh of the same connectivity can copy them and compute only its values
// ---------------------------------------------------------------------------
template <class T> static void clone_buf(afe_ctx *ctx, DevBuf<T> &dst, const DevBuf<T> &src) {
  if (!src.p || src.n == 0) {
    dst.release();
    return;
  }
  dst.alloc(ctx, src.n);
  HIPCHK(hipMemcpyAsync(dst.p, src.p, src.n * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream));
}

// a copy of src (another mesh of the same connectivity) owned by context
// ctx, on src's device: its device arrays are copied on ctx's stream, so src
// must not be written meanwhile
static int space_clone_into(afe_ctx *ctx, const afe_space *src, const double *coords, afe_space **out) {
  API_BEGIN
  if (!src || !out || !ctx) return fail(AFE_E_ARG, "null argument");
  if (!src->ctx) return fail(AFE_E_ARG, "host-only space");
  if (src->ctx->device != ctx->device) return fail(AFE_E_ARG, "space of another device");
  auto s = std::make_unique<afe_space>();
  s->ctx = ctx;
  s->h = src->h;
  if (coords) std::memcpy(s->h.coords.data(), coords, s->h.coords.size() * sizeof(double));
  s->seg_max = src->seg_max;
  s->gather_dep = src->gather_dep;
  s->distinct_dofs = src->distinct_dofs;
  std::lock_guard<CtxMutex> lk(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  s->coords.upload(ctx, s->h.coords);
  clone_buf(ctx, s->ev, src->ev);
  clone_buf(ctx, s->elem_dof, src->elem_dof);
  clone_buf(ctx, s->dof_mark, src->dof_mark);
  clone_buf(ctx, s->adj, src->adj);
  clone_buf(ctx, s->col, src->col);
  clone_buf(ctx, s->adj_ptr, src->adj_ptr);
  clone_buf(ctx, s->rp, src->rp);
  clone_buf(ctx, s->slot, src->slot);
  HIPCHK(hipStreamSynchronize(ctx->stream));
  *out = s.release();
  API_END
}

extern "C" int afe_space_clone(const afe_space *src, const double *coords, afe_space **out) {
  if (!src) return fail(AFE_E_ARG, "null argument");
  return space_clone_into(src->ctx, src, coords, out);
}

extern "C" int afe_amg_clone(const afe_amg *src, afe_matrix *a, afe_amg **out) {
  API_BEGIN
  if (!src || !a || !out || !src->ctx) return fail(AFE_E_ARG, "null argument");
  if (!src->prm.lazy || !src->dev_refresh)
    return fail(AFE_E_ARG, "only a lazy hierarchy with device refresh maps can be cloned");
  if (!a->ctx || a->ctx->device != src->ctx->device) return fail(AFE_E_ARG, "matrix of another device");
  const Csr &F = src->h.A[0];
  if (a->m != F.m || a->nnz != F.nnz() || a->hrp() != F.rp || a->hcol() != F.col)
    return fail(AFE_E_PATTERN, "the matrix does not have the hierarchy's sparsity pattern");
  // the copy belongs to the matrix's context (the source's may be another
  // one on the same device: its buffers are read on this stream)
  afe_ctx *ctx = a->ctx;
  auto g = std::make_unique<afe_amg>();
  g->ctx = ctx;
  g->A0 = a;
  g->key_space = a->s;
  g->key_mat = a->s ? nullptr : a;
  g->prm = src->prm;
  g->h = src->h; // host copy before the context lock: other callers keep the device meanwhile
  std::lock_guard<CtxMutex> lk(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  g->has32 = src->has32;
  g->tail0 = src->tail0;
  g->coarse_exact = src->coarse_exact;
  g->coarse_mode = src->coarse_mode;
  g->big_coarse = src->big_coarse;
  g->dev_refresh = true;
  g->nat_ns = src->nat_ns;
  g->graph_epoch = g_tune.epoch;
  g->lv.resize(src->lv.size());
  for (size_t l = 0; l < src->lv.size(); l++) {
    DevLevel &d = g->lv[l];
    const DevLevel &o = src->lv[l];
    d.n = o.n;
    d.ncol = o.ncol;
    d.colour_slice = o.colour_slice;
    d.rows_c0 = o.rows_c0;
    d.colour_row = o.colour_row;
    d.ns = o.ns;
    d.p_ns = o.p_ns;
    d.t_ns = o.t_ns;
    d.gs = o.gs;
    d.p_gs = o.p_gs;
    d.zs = o.zs;
    d.cshift = o.cshift;
    d.perm = o.perm;
    d.stream = o.stream;
    clone_buf(ctx, d.si, o.si);
    clone_buf(ctx, d.p_si, o.p_si);
    clone_buf(ctx, d.t_si, o.t_si);
    clone_buf(ctx, d.cols, o.cols);
    clone_buf(ctx, d.p_cols, o.p_cols);
    clone_buf(ctx, d.t_cols, o.t_cols);
    clone_buf(ctx, d.col_slice, o.col_slice);
    clone_buf(ctx, d.c16, o.c16);
    clone_buf(ctx, d.cb16, o.cb16);
    // values too: the refresh below rewrites every stored entry, and the
    // copies bring the zero padding of the sliced-ELL slices with them
    clone_buf(ctx, d.v64, o.v64);
    clone_buf(ctx, d.invd64, o.invd64);
    clone_buf(ctx, d.diag64, o.diag64);
    clone_buf(ctx, d.p_v64, o.p_v64);
    clone_buf(ctx, d.t_v64, o.t_v64);
    clone_buf(ctx, d.v32, o.v32);
    clone_buf(ctx, d.invd32, o.invd32);
    clone_buf(ctx, d.diag32, o.diag32);
    clone_buf(ctx, d.p_v32, o.p_v32);
    clone_buf(ctx, d.t_v32, o.t_v32);
    clone_buf(ctx, d.c_rp, o.c_rp);
    clone_buf(ctx, d.P_rp, o.P_rp);
    clone_buf(ctx, d.Pt_rp, o.Pt_rp);
    clone_buf(ctx, d.c_col, o.c_col);
    clone_buf(ctx, d.P_col, o.P_col);
    clone_buf(ctx, d.Pt_col, o.Pt_col);
    clone_buf(ctx, d.smap, o.smap);
    clone_buf(ctx, d.dsrc, o.dsrc);
    clone_buf(ctx, d.c_val, o.c_val);
    clone_buf(ctx, d.P_val, o.P_val);
    clone_buf(ctx, d.Pt_val, o.Pt_val);
    d.x64.alloc(ctx, d.n);
    d.b64.alloc(ctx, d.n);
    d.t64.alloc(ctx, d.n);
    if (g->has32) {
      d.x32.alloc(ctx, d.n);
      d.b32.alloc(ctx, d.n);
      d.t32.alloc(ctx, d.n);
    }
  }
  clone_buf(ctx, g->perm0, src->perm0);
  clone_buf(ctx, g->iperm0, src->iperm0);
  clone_buf(ctx, g->nat_si, src->nat_si);
  clone_buf(ctx, g->nat_cols, src->nat_cols);
  clone_buf(ctx, g->nat_rowmap, src->nat_rowmap);
  clone_buf(ctx, g->nat_smap, src->nat_smap);
  clone_buf(ctx, g->nat_dsrc, src->nat_dsrc);
  clone_buf(ctx, g->nat_c16, src->nat_c16);
  clone_buf(ctx, g->nat_cbase, src->nat_cbase);
  clone_buf(ctx, g->nat_v64, src->nat_v64);
  clone_buf(ctx, g->nat_diag, src->nat_diag);
  clone_buf(ctx, g->minv, src->minv);
  HIPCHK(hipHostMalloc((void **)&g->h_sc, sizeof(CgScalars), hipHostMallocDefault));
  finish_levels(g.get());
  device_lazy_refresh(g.get(), a);
  *out = g.release();
  API_END
}

extern "C" int afe_amg_info(const afe_amg *g, int32_t *n_levels, int64_t *level_n, int64_t *level_nnz,
                            int32_t *coarse_full) {
  if (!g) return fail(AFE_E_ARG, "null argument");
  const int L = (int)g->h.A.size();
  if (n_levels) *n_levels = L;
  for (int l = 0; l < L; l++) {
    if (level_n) level_n[l] = g->h.A[l].m;
    if (level_nnz) level_nnz[l] = g->h.A[l].nnz();
  }
  if (coarse_full) *coarse_full = g->h.full;
  return AFE_OK;
}

// colour of every row of a device level, in the level's own row numbering
extern "C" int afe_amg_level_colours(const afe_amg *g, int32_t level, int32_t *colour) {
  if (!g || !colour) return fail(AFE_E_ARG, "null argument");
  if (level < 0 || level >= (int)g->lv.size()) return fail(AFE_E_ARG, "no such level");
  const DevLevel &d = g->lv[level];
  for (int c = 0; c + 1 < (int)d.colour_row.size(); c++)
    for (int32_t r = d.colour_row[c]; r < d.colour_row[c + 1]; r++) colour[d.perm[r]] = c;
  return AFE_OK;
}

extern "C" int afe_amg_level_csr(const afe_amg *g, int32_t level, int32_t which, int64_t *m, int64_t *n,
                                 int64_t *nnz, int64_t *row_ptr, int64_t *col, double *val) {
  if (!g) return fail(AFE_E_ARG, "null argument");
  const int L = (int)g->h.A.size();
  if (level < 0 || level >= L || (which == 1 && level >= L - 1) || which < 0 || which > 1)
    return fail(AFE_E_ARG, "no such level matrix");
  if (val && which == 0 && g->ctx) {
    std::lock_guard<CtxMutex> lk(g->ctx->mu);
    try {
      sync_host_values(const_cast<afe_amg *>(g));
    } catch (const HipFail &h) {
      return fail(AFE_E_HIP, std::string(h.what) + ": " + hipGetErrorString(h.e));
    }
  }
  const Csr &c = which == 0 ? g->h.A[level] : g->h.P[level];
  if (m) *m = c.m;
  if (n) *n = c.n;
  if (nnz) *nnz = c.nnz();
  if (row_ptr) std::memcpy(row_ptr, c.rp.data(), sizeof(int64_t) * (c.m + 1));
  if (col)
    for (int64_t k = 0; k < c.nnz(); k++) col[k] = c.col[k];
  if (val) std::memcpy(val, c.val.data(), sizeof(double) * c.nnz());
  return AFE_OK;
}

// ---------------------------------------------------------------------------
// V-cycle issue (capturable: kernels only)
// ---------------------------------------------------------------------------
template <class T> struct Prec;
template <> struct Prec<double> {
  static const double *v(const DevLevel &d) { return d.v64.p; }
  static const double *invd(const DevLevel &d) { return d.invd64.p; }
  static const double *diag(const DevLevel &d) { return d.diag64.p; }
  static const double *pv(const DevLevel &d) { return d.p_v64.p; }
  static const double *tv(const DevLevel &d) { return d.t_v64.p; }
  static double *x(DevLevel &d) { return d.x64.p; }
  static double *b(DevLevel &d) { return d.b64.p; }
  static double *t(DevLevel &d) { return d.t64.p; }
  static const LevelDev<double> *ld(afe_amg *g) { return g->ld64.p; }
};
template <> struct Prec<float> {
  static const float *v(const DevLevel &d) { return d.v32.p; }
  static const float *invd(const DevLevel &d) { return d.invd32.p; }
  static const float *diag(const DevLevel &d) { return d.diag32.p; }
  static const float *pv(const DevLevel &d) { return d.p_v32.p; }
  static const float *tv(const DevLevel &d) { return d.t_v32.p; }
  static float *x(DevLevel &d) { return d.x32.p; }
  static float *b(DevLevel &d) { return d.b32.p; }
  static float *t(DevLevel &d) { return d.t32.p; }
  static const LevelDev<float> *ld(afe_amg *g) { return g->ld32.p; }
};

template <class T>
static void gs_colour(DevLevel &d, int c, hipStream_t st) {
  const int s0 = d.colour_slice[c], s1 = d.colour_slice[c + 1];
  if (s1 <= s0) return;
  if (d.stream)
    k_gs<T, true><<<sblk(s1 - s0), BLK, 0, st>>>(d.si.p, s0, s1, d.cols.p, d.c16.p, d.cb16.p, d.cshift,
                                                 Prec<T>::v(d), Prec<T>::invd(d),
                                                 Prec<T>::b(d), Prec<T>::x(d), d.gs, g_tune.xcd);
  else
    k_gs<T, false><<<sblk(s1 - s0), BLK, 0, st>>>(d.si.p, s0, s1, d.cols.p, nullptr, nullptr, 0,
                                                  Prec<T>::v(d), Prec<T>::invd(d),
                                                  Prec<T>::b(d), Prec<T>::x(d), d.gs, g_tune.xcd);
}

// one V(nu, nu) cycle on level-0 vectors (b, x) of precision T; the
// post-smoother runs the colours in reverse order when symmetric.  x0_zero:
// the cycle starts from x = 0 on level 0, and the first colour kernel writes
// all of level 0's x (the caller need not clear it).
template <class T>
static void issue_vcycle(afe_amg *g, int nu, bool symmetric, hipStream_t st, bool x0_zero) {
  const int Lc = (int)g->lv.size() - 1;
  for (int l = 0; l < g->tail0; l++) {
    DevLevel &d = g->lv[l];
    DevLevel &dn = g->lv[l + 1];
    if (l == 0 && x0_zero && nu == 0) k_fill<T><<<nblk(d.n), 256, 0, st>>>(d.n, Prec<T>::x(d), (T)0);
    // the next level is smoothed by launches from zero: the restriction
    // takes its first colour step
    const bool first_next = nu > 0 && l + 1 < g->tail0;
    for (int it = 0; it < nu; it++)
      for (int c = 0; c < d.ncol; c++) {
        if (l == 0 && it == 0 && c == 0 && x0_zero) {
          k_gs_first<T><<<nblk(d.n), 256, 0, st>>>(d.n, d.rows_c0, Prec<T>::b(d), Prec<T>::invd(d), Prec<T>::x(d));
          continue;
        }
        if (l > 0 && it == 0 && c == 0) continue; // done by the restriction from level l - 1
        gs_colour<T>(d, c, st);
      }
    if (d.stream) {
      k_residual<T, true><<<sblk(d.ns), BLK, 0, st>>>(d.si.p, d.ns, nu > 0 ? d.zs : d.ns, d.cols.p, d.c16.p,
                                                       d.cb16.p, d.cshift,
                                                       Prec<T>::v(d), Prec<T>::diag(d),
                                                       Prec<T>::b(d), Prec<T>::x(d), Prec<T>::t(d), d.gs,
                                                       g_tune.xcd);
      k_restrict<T, true><<<sblk(d.p_ns), BLK, 0, st>>>(d.p_si.p, d.p_ns, d.p_cols.p, Prec<T>::pv(d),
                                                         Prec<T>::t(d), Prec<T>::b(dn), Prec<T>::x(dn),
                                                         first_next ? Prec<T>::invd(dn) : nullptr, (int)dn.rows_c0,
                                                         d.p_gs, g_tune.xcd);
    } else {
      k_residual<T, false><<<sblk(d.ns), BLK, 0, st>>>(d.si.p, d.ns, nu > 0 ? d.zs : d.ns, d.cols.p, nullptr,
                                                        nullptr, 0,
                                                        Prec<T>::v(d), Prec<T>::diag(d),
                                                        Prec<T>::b(d), Prec<T>::x(d), Prec<T>::t(d), d.gs,
                                                        g_tune.xcd);
      k_restrict<T, false><<<sblk(d.p_ns), BLK, 0, st>>>(d.p_si.p, d.p_ns, d.p_cols.p, Prec<T>::pv(d),
                                                          Prec<T>::t(d), Prec<T>::b(dn), Prec<T>::x(dn),
                                                          first_next ? Prec<T>::invd(dn) : nullptr,
                                                          (int)dn.rows_c0, d.p_gs, g_tune.xcd);
    }
  }
  const int lds_cap = (int)g->ctx->lds_bytes; // dynamic LDS of the tail workgroup
  const size_t lds = (size_t)lds_cap;
  if (g->big_coarse) {
    DevLevel &dc = g->lv[Lc];
    if (g->tail0 < Lc)
      k_tail<T><<<1, TAIL_THREADS, lds, st>>>(Prec<T>::ld(g), g->tail0, Lc, g->minv.p, g->coarse_mode, nu, nu,
                                               symmetric ? 1 : 0, 1, lds_cap);
    k_coarse_gemv<T><<<nblk(dc.n, WPB), BLK, 0, st>>>((int)dc.n, g->minv.p, Prec<T>::b(dc), Prec<T>::x(dc));
    if (g->tail0 < Lc)
      k_tail<T><<<1, TAIL_THREADS, lds, st>>>(Prec<T>::ld(g), g->tail0, Lc, g->minv.p, g->coarse_mode, nu, nu,
                                               symmetric ? 1 : 0, 4, lds_cap);
  } else {
    k_tail<T><<<1, TAIL_THREADS, lds, st>>>(Prec<T>::ld(g), g->tail0, Lc, g->minv.p, g->coarse_mode, nu, nu,
                                             symmetric ? 1 : 0, 7, lds_cap);
  }
  for (int l = g->tail0 - 1; l >= 0; l--) {
    DevLevel &d = g->lv[l];
    DevLevel &dn = g->lv[l + 1];
    // the colour the post-smoother sweeps first is left to it
    const int lc0 = d.ncol > 0 ? d.colour_row[d.ncol - 1] : (int)d.n;
    const int skip0 = nu == 0 ? 0 : (symmetric ? lc0 : 0);
    const int skip1 = nu == 0 ? 0 : (symmetric ? (int)d.n : (int)d.rows_c0);
    if (d.stream)
      k_prolong<T, true><<<sblk(d.t_ns), BLK, 0, st>>>(d.t_si.p, d.t_ns, d.t_cols.p, Prec<T>::tv(d),
                                                        Prec<T>::x(dn), Prec<T>::x(d), skip0, skip1, g_tune.xcd);
    else
      k_prolong<T, false><<<sblk(d.t_ns), BLK, 0, st>>>(d.t_si.p, d.t_ns, d.t_cols.p, Prec<T>::tv(d),
                                                         Prec<T>::x(dn), Prec<T>::x(d), skip0, skip1, g_tune.xcd);
    for (int it = 0; it < nu; it++)
      for (int cc = 0; cc < d.ncol; cc++) gs_colour<T>(d, symmetric ? d.ncol - 1 - cc : cc, st);
  }
}

// captures what issue() puts on stream st into an executable graph.  If
// issue() throws, the capture is ended and its partial graph destroyed
// before the error propagates, so the stream never stays in capture mode.
template <class F> static hipGraphExec_t capture_graph(hipStream_t st, F issue) {
  HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  try {
    issue();
  } catch (...) {
    hipGraph_t partial = nullptr;
    if (hipStreamEndCapture(st, &partial) == hipSuccess && partial) (void)hipGraphDestroy(partial);
    throw;
  }
  hipGraph_t graph = nullptr;
  HIPCHK(hipStreamEndCapture(st, &graph));
  hipGraphExec_t ex = nullptr;
  const hipError_t e = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  HIPCHK(e);
  return ex;
}

// AFE_NO_GRAPH=1 issues the captured sequences as plain launches (for tools
// that do not follow graph launches)
static bool use_graphs() {
  static const bool v = getenv("AFE_NO_GRAPH") == nullptr;
  return v;
}

static void launch_graph_or_issue(afe_amg *g, int prec, bool symmetric, int nu) {
  hipStream_t st = g->ctx->stream;
  if (!use_graphs()) {
    if (prec == 1)
      issue_vcycle<float>(g, nu, symmetric, st, symmetric);
    else
      issue_vcycle<double>(g, nu, symmetric, st, symmetric);
    return;
  }
  if (g->graph_epoch != g_tune.epoch) { // launch shapes changed since capture
    drop_graphs(g);
    g->graph_epoch = g_tune.epoch;
  }
  hipGraphExec_t &ex = g->vc_graph[prec][symmetric ? 1 : 0];
  if (ex && g->graph_nu[prec][symmetric ? 1 : 0] != nu) {
    (void)hipGraphExecDestroy(ex);
    ex = nullptr;
  }
  if (!ex) {
    ex = capture_graph(st, [&]() {
      if (prec == 1)
        issue_vcycle<float>(g, nu, symmetric, st, symmetric);
      else
        issue_vcycle<double>(g, nu, symmetric, st, symmetric);
    });
    g->graph_nu[prec][symmetric ? 1 : 0] = nu;
  }
  HIPCHK(hipGraphLaunch(ex, st));
}


// true residual r = b - A (x + xl) in compensated arithmetic, and its norms
// natural = true: vectors in the natural numbering (the CG), else in the
// level-0 colour order (the stationary solver)
static void residual_norms(afe_amg *g, bool natural, const double *b, const double *x, double *r,
                           const double *xl = nullptr, hipStream_t stream = nullptr) {
  DevLevel &d = g->lv[0];
  hipStream_t st = stream ? stream : g->ctx->stream;
  const int G = g_tune.spmv_blocks;
  if (natural)
    k_residual_norms_dd<<<G, BLK, 0, st>>>(g->nat_si.p, g->nat_ns, g->nat_cols.p, g->nat_c16.p, g->nat_cbase.p,
                                            g->nat_v64.p, g->nat_diag.p, g->nat_rowmap.p, b, x, xl, r, g->part.p,
                                            g_tune.xcd);
  else
    k_residual_norms_dd<<<G, BLK, 0, st>>>(d.si.p, d.ns, d.cols.p, nullptr, nullptr, d.v64.p, d.diag64.p, nullptr,
                                            b, x, xl, r, g->part.p, g_tune.xcd);
  k_finalize<<<1, FIN_THREADS, 0, st>>>(g->part.p, G, 3, &g->sc.p->rr);
}

static void read_scalars(afe_amg *g) {
  HIPCHK(hipMemcpyAsync(g->h_sc, g->sc.p, sizeof(CgScalars), hipMemcpyDeviceToHost, g->ctx->stream));
  HIPCHK(hipStreamSynchronize(g->ctx->stream));
}

extern "C" int afe_amg_solve(afe_amg *g, double *d_x, const double *d_r, double tol, int64_t step, int32_t mode,
                             double rtol, afe_solve_stats *st) {
  API_BEGIN
  if (!g || !d_x || !d_r || step < 0) return fail(AFE_E_ARG, "bad solve arguments");
  std::lock_guard<CtxMutex> lk(g->ctx->mu);
  HIPCHK(hipSetDevice(g->ctx->device));
  afe_ctx *ctx = g->ctx;
  hipStream_t s = ctx->stream;
  DevLevel &d0 = g->lv[0];
  const int64_t n = d0.n;
  if (tol == 0.0) tol = 1.0e-12; // AMGSolver::tolerence() default
  HIPCHK(hipEventRecord(ctx->ev0, s));
  k_permute_in<double, double><<<nblk(n), 256, 0, s>>>(n, g->perm0.p, d_r, d0.b64.p);
  k_permute_in<double, double><<<nblk(n), 256, 0, s>>>(n, g->perm0.p, d_x, d0.x64.p);
  residual_norms(g, false, d0.b64.p, d0.x64.p, nullptr);
  read_scalars(g);
  const double init_l1 = g->h_sc->l1, bb = g->h_sc->bb;
  double rr = g->h_sc->rr, l1 = init_l1;
  int64_t it = 0;
  bool conv = false;
  if (rtol <= 0 && !(mode & 2) && !verbose()) {
    // the reference's form (no early exit, nothing printed per cycle): the
    // step + 1 cycles run back to back, the residual is taken once after them
    for (; it <= step; it++) launch_graph_or_issue(g, 0, false, g->prm.smooth);
    residual_norms(g, false, d0.b64.p, d0.x64.p, nullptr);
    read_scalars(g);
    rr = g->h_sc->rr;
    l1 = g->h_sc->l1;
  } else {
    while (true) {
      launch_graph_or_issue(g, 0, false, g->prm.smooth);
      residual_norms(g, false, d0.b64.p, d0.x64.p, nullptr);
      read_scalars(g);
      rr = g->h_sc->rr;
      l1 = g->h_sc->l1;
      it++;
      if (verbose()) fprintf(stderr, "cycle %lld rel res %.3e l1 %.3e\n", (long long)it, std::sqrt(rr / bb), l1);
      if (rtol > 0 && rr <= rtol * rtol * bb) {
        conv = true;
        break;
      }
      if (it > step) break;
      if (mode & 2) fprintf(stderr, "\r\tresidual = %g", l1);
    }
  }
  k_permute_out<double, double><<<nblk(n), 256, 0, s>>>(n, g->perm0.p, d0.x64.p, d_x);
  HIPCHK(hipEventRecord(ctx->ev1, s));
  HIPCHK(hipEventSynchronize(ctx->ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  if (rtol <= 0) conv = l1 < tol * init_l1;
  if ((mode & 1) == 0)
    fprintf(stderr, "\r\t%s with residual %g at step %lld.\n", conv ? "converge" : "failed to converge", l1,
            (long long)it);
  if (st) {
    st->iterations = (int32_t)it;
    st->converged = conv ? 1 : 0;
    st->rel_res_l2 = bb > 0 ? std::sqrt(rr / bb) : 0.0;
    st->res_l1 = l1;
    st->rel_res_l1 = init_l1 > 0 ? l1 / init_l1 : 0.0;
    st->seconds = ms * 1e-3;
  }
  API_END
}

extern "C" int afe_amg_precondition(afe_amg *g, double *d_x, const double *d_b, int32_t iterate_step) {
  afe_solve_stats st;
  return afe_amg_solve(g, d_x, d_b, 0.0, iterate_step, 1, 0.0, &st);
}

// z = M^{-1} r with one symmetric V-cycle from zero; r and z in the natural
// numbering, the V-cycle in the level-0 colour order
template <class T> static void precondition_perm(afe_amg *g, const double *r, bool first) {
  DevLevel &d0 = g->lv[0];
  hipStream_t s = g->ctx->stream;
  const int64_t n = d0.n;
  k_prep_vcycle<T><<<nblk(n), 256, 0, s>>>(n, r, Prec<T>::b(d0), Prec<T>::x(d0), g->iperm0.p);
  launch_graph_or_issue(g, std::is_same<T, float>::value ? 1 : 0, true, g->prm.smooth);
  if (first) k_fill<double><<<nblk(n), 256, 0, s>>>(n, g->wz.p, 0.0);
  k_take_z<T><<<RED_BLOCKS, 256, 0, s>>>(n, Prec<T>::x(d0), r, g->wz.p, g->iperm0.p, g->part.p);
  k_finalize_beta<<<1, FIN_THREADS, 0, s>>>(g->part.p, RED_BLOCKS, g->sc.p, first ? 1 : 0);
  k_p_update<<<RED_BLOCKS, 256, 0, s>>>(n, g->sc.p, g->wz.p, g->wp.p);
}

// one CG iteration after the direction p is set: q = A p, alpha, x/r update,
// rr to the host slot, then the next preconditioned direction
// q = A p and (p, q)
// with start/stop events the two timestamps are those of the kernel's own
// dispatch (hipExtLaunchKernel), so host launch gaps stay out of the figure
static void issue_cg_spmv(afe_amg *g, hipStream_t s, hipEvent_t start = nullptr, hipEvent_t stop = nullptr) {
  if (start) {
    hipExtLaunchKernelGGL(k_spmv_dot, dim3(g_tune.spmv_blocks), dim3(BLK), 0, s, start, stop, 0,
                          g->nat_si.p, g->nat_ns, g->nat_cols.p, g->nat_c16.p, g->nat_cbase.p, g->nat_v64.p,
                          g->nat_diag.p, g->nat_rowmap.p, g->wp.p, g->wq.p, g->part.p, g_tune.xcd);
    HIPCHK(hipGetLastError());
    return;
  }
  k_spmv_dot<<<g_tune.spmv_blocks, BLK, 0, s>>>(g->nat_si.p, g->nat_ns, g->nat_cols.p, g->nat_c16.p,
                                                 g->nat_cbase.p, g->nat_v64.p, g->nat_diag.p, g->nat_rowmap.p,
                                                 g->wp.p, g->wq.p, g->part.p, g_tune.xcd);
}

// part 0 of a CG iteration after the SpMV: alpha, the x / r update with the
// FP32 (or FP64) right-hand side of the V-cycle, (r, r), and the scalars to
// the host
template <class T> static void issue_cg_update(afe_amg *g) {
  DevLevel &d0 = g->lv[0];
  hipStream_t s = g->ctx->stream;
  const int64_t n = d0.n;
  k_finalize<<<1, FIN_THREADS, 0, s>>>(g->part.p, g_tune.spmv_blocks, 1, &g->sc.p->pq);
  // level 0 runs as launches (tail0 > 0): the V-cycle's first colour kernel
  // sets x0, so the update does not clear it
  k_cg_update<T><<<RED_BLOCKS, 256, 0, s>>>(n, g->sc.p, g->wx.p, g->wxl.p, g->wr.p, g->wp.p, g->wq.p,
                                             Prec<T>::b(d0), g->tail0 > 0 ? nullptr : Prec<T>::x(d0),
                                             g->iperm0.p, g->part.p);
  k_finalize_host<<<1, FIN_THREADS, 0, s>>>(g->part.p, RED_BLOCKS, &g->sc.p->rr, g->sc.p, g->h_sc);
}

// part 1: z = M r (the V-cycle, issued inline: nested capture is not used),
// the flexible beta and p = z + beta p
template <class T> static void issue_cg_precondition(afe_amg *g) {
  DevLevel &d0 = g->lv[0];
  hipStream_t s = g->ctx->stream;
  const int64_t n = d0.n;
  issue_vcycle<T>(g, g->prm.smooth, true, s, true);
  k_take_z<T><<<RED_BLOCKS, 256, 0, s>>>(n, Prec<T>::x(d0), g->wr.p, g->wz.p, g->iperm0.p, g->part.p);
  k_finalize_beta<<<1, FIN_THREADS, 0, s>>>(g->part.p, RED_BLOCKS, g->sc.p, 0);
  k_p_update<<<RED_BLOCKS, 256, 0, s>>>(n, g->sc.p, g->wz.p, g->wp.p);
}

// launches one captured part (capturing it on first use or after a change
// of the smoothing count)
template <class F> static void launch_part(afe_amg *g, int part, int prec, F issue) {
  hipStream_t st = g->ctx->stream;
  if (!use_graphs()) {
    issue();
    return;
  }
  if (g->graph_epoch != g_tune.epoch) {
    drop_graphs(g);
    g->graph_epoch = g_tune.epoch;
  }
  hipGraphExec_t &ex = g->cg_graph[part][prec];
  if (ex && g->cg_nu[part][prec] != g->prm.smooth) {
    (void)hipGraphExecDestroy(ex);
    ex = nullptr;
  }
  if (!ex) {
    ex = capture_graph(st, issue);
    g->cg_nu[part][prec] = g->prm.smooth;
  }
  HIPCHK(hipGraphLaunch(ex, st));
}

// one CG iteration: part 0 (with the SpMV; with live timing the SpMV is
// launched directly with its dispatch-timed events), the event it_ev, part 1.
// The host waits for it_ev only: it reads (r, r) and queues the next
// iteration while part 1 still runs, so the device does not idle between
// iterations.
template <class T> static void launch_cg_iteration(afe_amg *g) {
  const int prec = std::is_same<T, float>::value ? 1 : 0;
  hipStream_t st = g->ctx->stream;
  if (!g->it_ev) HIPCHK(hipEventCreateWithFlags(&g->it_ev, hipEventDisableTiming));
  hipEvent_t *pe = g->parts ? g->pev[g->part_set] : nullptr;
  if (pe) HIPCHK(hipEventRecord(pe[0], st));
  if (g->live) issue_cg_spmv(g, st, g->live0, g->live1);
  launch_part(g, 0, prec, [&]() {
    if (!g->live) issue_cg_spmv(g, st);
    issue_cg_update<T>(g);
  });
  HIPCHK(hipEventRecord(g->it_ev, st));
  if (pe) HIPCHK(hipEventRecord(pe[1], st));
  launch_part(g, 1, prec, [&]() { issue_cg_precondition<T>(g); });
  if (pe) HIPCHK(hipEventRecord(pe[2], st));
}

// adds the part times of the iteration recorded in event set `set` (its
// events have completed)
static void part_read(afe_amg *g, int set) {
  float a = 0, b = 0;
  if (hipEventElapsedTime(&a, g->pev[set][0], g->pev[set][1]) == hipSuccess &&
      hipEventElapsedTime(&b, g->pev[set][1], g->pev[set][2]) == hipSuccess) {
    g->part_ms[0] += a;
    g->part_ms[1] += b;
    g->part_n++;
  } else {
    (void)hipGetLastError();
  }
}

// after the host's wait for an iteration's part 0: the previous iteration's
// set is complete (its part 1 preceded this part 0 on the stream)
static void part_step(afe_amg *g) {
  if (!g->parts) return;
  if (g->part_pending) part_read(g, g->part_set ^ 1);
  g->part_pending = true;
  g->part_set ^= 1;
}

template <class T>
static int pcg_impl(afe_amg *g, double *d_x, const double *d_b, double rtol, int max_it, afe_solve_stats *st) {
  afe_ctx *ctx = g->ctx;
  hipStream_t s = ctx->stream;
  DevLevel &d0 = g->lv[0];
  const int64_t n = d0.n;
  HIPCHK(hipEventRecord(ctx->ev0, s));
  HIPCHK(hipMemcpyAsync(g->wb.p, d_b, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(g->wx.p, d_x, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  k_fill<double><<<nblk(n), 256, 0, s>>>(n, g->wxl.p, 0.0);
  residual_norms(g, true, g->wb.p, g->wx.p, g->wr.p, g->wxl.p);
  read_scalars(g);
  const double bb = g->h_sc->bb;
  const double init_l1 = g->h_sc->l1;
  double rr = g->h_sc->rr, l1 = init_l1;
  const double thr = rtol * rtol * bb;
  // the pair x_hi + x_lo is driven to rtol/2 so that the single rounding to
  // the delivered FP64 vector keeps its true residual below rtol
  double pthr = 0.25 * thr;
  int it = 0;
  bool conv = rr <= pthr;
  int restarts = 0;
  for (int outer = 0; outer < 4; outer++) {
  while (!conv && it < max_it) {
    k_fill<double><<<nblk(n), 256, 0, s>>>(n, g->wp.p, 0.0);
    precondition_perm<T>(g, g->wr.p, true);
    while (it < max_it) {
      launch_cg_iteration<T>(g);
      HIPCHK(hipEventSynchronize(g->it_ev));
      part_step(g);
      if (g->live) {
        // the pair is recorded by the kernel's own dispatch; a tool that
        // replaces the dispatch's completion signal leaves it unrecorded,
        // and that launch is then counted as untimed instead of failing
        float ms = 0;
        if (hipEventElapsedTime(&ms, g->live0, g->live1) == hipSuccess) {
          g->live_ms += ms;
          g->live_n++;
        } else {
          (void)hipGetLastError();
          g->live_missed++;
        }
      }
      it++;
      rr = g->h_sc->rr;
      if (verbose())
        fprintf(stderr, "pcg it %d rel res %.3e pq %.3e rz %.3e beta %.3e\n", it, std::sqrt(rr / bb),
                g->h_sc->pq, g->h_sc->rz, g->h_sc->beta);
      if (!(rr == rr)) {
        HIPCHK(hipStreamSynchronize(s)); // part 1 of the iteration is still queued
        g->part_pending = false;
        return fail(AFE_E_NOTCONVERGED, "CG breakdown (NaN residual)");
      }
      if (rr <= pthr) break;
    }
    // true residual of the pair x_hi + x_lo; restart from it when the
    // recurrence drifted
    residual_norms(g, true, g->wb.p, g->wx.p, g->wr.p, g->wxl.p);
    read_scalars(g);
    rr = g->h_sc->rr;
    conv = rr <= pthr;
    if (!conv && ++restarts > 20) break;
  }
  // deliver round(x_hi + x_lo) and check that vector's true residual
  k_dd_round<<<nblk(n), 256, 0, s>>>(n, g->wx.p, g->wxl.p);
  residual_norms(g, true, g->wb.p, g->wx.p, g->wr.p, nullptr);
  read_scalars(g);
  rr = g->h_sc->rr;
  l1 = g->h_sc->l1;
  if (verbose())
    fprintf(stderr, "pcg delivered vector: rel res %.3e (pair converged %d)\n", std::sqrt(rr / bb), conv ? 1 : 0);
  conv = rr <= thr;
  if (conv || it >= max_it || restarts > 20) break;
  pthr *= 0.25; // rounding floor reached: tighten the pair and continue
  }
  HIPCHK(hipMemcpyAsync(d_x, g->wx.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  HIPCHK(hipEventRecord(ctx->ev1, s));
  HIPCHK(hipEventSynchronize(ctx->ev1));
  if (g->parts && g->part_pending) { // the last iteration's set
    part_read(g, g->part_set ^ 1);
    g->part_pending = false;
  }
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  if (st) {
    st->iterations = it;
    st->converged = conv ? 1 : 0;
    st->rel_res_l2 = bb > 0 ? std::sqrt(rr / bb) : 0.0;
    st->res_l1 = l1;
    st->rel_res_l1 = init_l1 > 0 ? l1 / init_l1 : 0.0;
    st->seconds = ms * 1e-3;
  }
  if (!conv) return fail(AFE_E_NOTCONVERGED, "PCG did not reach the tolerance");
  return AFE_OK;
}

extern "C" int afe_amg_pcg(afe_amg *g, double *d_x, const double *d_b, double rtol, int32_t max_it, int32_t prec,
                           afe_solve_stats *st) {
  API_BEGIN
  if (!g || !d_x || !d_b || rtol <= 0 || max_it < 1) return fail(AFE_E_ARG, "bad PCG arguments");
  if (prec == AFE_PREC_FP32 && !g->has32) return fail(AFE_E_ARG, "hierarchy built without FP32 copies");
  std::lock_guard<CtxMutex> lk(g->ctx->mu);
  HIPCHK(hipSetDevice(g->ctx->device));
  return prec == AFE_PREC_FP32 ? pcg_impl<float>(g, d_x, d_b, rtol, max_it, st)
                               : pcg_impl<double>(g, d_x, d_b, rtol, max_it, st);
  API_END
}

extern "C" int afe_amg_vcycle(afe_amg *g, double *d_z, const double *d_r, int32_t prec) {
  API_BEGIN
  if (!g || !d_z || !d_r) return fail(AFE_E_ARG, "null argument");
  if (prec == AFE_PREC_FP32 && !g->has32) return fail(AFE_E_ARG, "hierarchy built without FP32 copies");
  std::lock_guard<CtxMutex> lk(g->ctx->mu);
  HIPCHK(hipSetDevice(g->ctx->device));
  hipStream_t s = g->ctx->stream;
  DevLevel &d0 = g->lv[0];
  const int64_t n = d0.n;
  k_permute_in<double, double><<<nblk(n), 256, 0, s>>>(n, g->perm0.p, d_r, g->wr.p);
  if (prec == AFE_PREC_FP32) {
    k_prep_vcycle<float><<<nblk(n), 256, 0, s>>>(n, g->wr.p, d0.b32.p, d0.x32.p, nullptr);
    launch_graph_or_issue(g, 1, true, g->prm.smooth);
    k_permute_out<float, double><<<nblk(n), 256, 0, s>>>(n, g->perm0.p, d0.x32.p, d_z);
  } else {
    k_prep_vcycle<double><<<nblk(n), 256, 0, s>>>(n, g->wr.p, d0.b64.p, d0.x64.p, nullptr);
    launch_graph_or_issue(g, 0, true, g->prm.smooth);
    k_permute_out<double, double><<<nblk(n), 256, 0, s>>>(n, g->perm0.p, d0.x64.p, d_z);
  }
  HIPCHK(hipStreamSynchronize(s));
  API_END
}

// ---------------------------------------------------------------------------
// timing of the level-0 kernels as launched by the solvers (HIP events on the
// context stream).  which: 0 FP64 SpMV with dot (CG q = A p), 1 one FP32
// Gauss-Seidel sweep over all colours, 2 one FP64 sweep, 3 FP32 residual.
// bytes: algorithmic bytes per launch (CSR formula of the reference pattern:
// nnz*(sv+4) + (m+1)*4 + n*sv + m*sv, sweeps count x read and written and r
// read: nnz*(sv+4) + (m+1)*4 + 3*m*sv).
// ---------------------------------------------------------------------------
extern "C" int afe_amg_kernel_timing(afe_amg *g, int32_t which, int32_t reps, double *avg_ms, double *bytes) {
  API_BEGIN
  if (!g || !g->ctx || reps < 1 || which < 0 || which > 3) return fail(AFE_E_ARG, "bad arguments");
  if ((which == 1 || which == 3) && !g->has32) return fail(AFE_E_ARG, "no FP32 copies");
  std::lock_guard<CtxMutex> lk(g->ctx->mu);
  HIPCHK(hipSetDevice(g->ctx->device));
  afe_ctx *ctx = g->ctx;
  hipStream_t s = ctx->stream;
  DevLevel &d = g->lv[0];
  const int64_t n = d.n, nnz = g->h.A[0].nnz();
  const double sv = (which == 1 || which == 3) ? 4.0 : 8.0;
  if (which == 0)
    *bytes = g->nat_c16.p ? nnz * (sv + 2) + (double)g->nat_ns * 4 + (n + 1) * 4.0 + 2.0 * n * sv
                          : nnz * (sv + 4) + (n + 1) * 4.0 + 2.0 * n * sv;
  else if (d.c16.p) // colour-coded columns: 2 bytes per entry and 8 bases per slice
    *bytes = nnz * (sv + 2) + (double)d.ns * CC_BASES * 4 + (n + 1) * 4.0 + 3.0 * n * sv;
  else
    *bytes = nnz * (sv + 4) + (n + 1) * 4.0 + 3.0 * n * sv;
  k_fill<double><<<nblk(n), 256, 0, s>>>(n, g->wp.p, 1.0);
  k_fill<double><<<nblk(n), 256, 0, s>>>(n, d.x64.p, 0.0);
  k_fill<double><<<nblk(n), 256, 0, s>>>(n, d.b64.p, 1.0);
  if (g->has32) {
    k_fill<float><<<nblk(n), 256, 0, s>>>(n, d.x32.p, 0.0f);
    k_fill<float><<<nblk(n), 256, 0, s>>>(n, d.b32.p, 1.0f);
  }
  auto once = [&]() {
    if (which == 0) {
      k_spmv_dot<<<g_tune.spmv_blocks, BLK, 0, s>>>(g->nat_si.p, g->nat_ns, g->nat_cols.p, g->nat_c16.p,
                                                     g->nat_cbase.p, g->nat_v64.p, g->nat_diag.p, g->nat_rowmap.p,
                                                     g->wp.p, g->wq.p, g->part.p, g_tune.xcd);
    } else if (which == 1) {
      for (int c = 0; c < d.ncol; c++) gs_colour<float>(d, c, s);
    } else if (which == 2) {
      for (int c = 0; c < d.ncol; c++) gs_colour<double>(d, c, s);
    } else {
      k_residual<float, true><<<sblk(d.ns), BLK, 0, s>>>(d.si.p, d.ns, d.ns, d.cols.p, d.c16.p, d.cb16.p, d.cshift,
                                                          d.v32.p, d.diag32.p, d.b32.p,
                                                          d.x32.p, d.t32.p, d.gs, g_tune.xcd);
    }
  };
  once();
  HIPCHK(hipEventRecord(ctx->ev0, s));
  for (int r = 0; r < reps; r++) once();
  HIPCHK(hipEventRecord(ctx->ev1, s));
  HIPCHK(hipEventSynchronize(ctx->ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  *avg_ms = ms / reps;
  API_END
}

// plain copy of `bytes` (rounded down to 16 B) between two scratch buffers,
// `reps` timed launches on the context stream
extern "C" int afe_copy_kernel_timing(afe_ctx *ctx, int64_t bytes, int32_t reps, double *avg_ms) {
  API_BEGIN
  if (!ctx || bytes < 16 || reps < 1 || !avg_ms) return fail(AFE_E_ARG, "bad arguments");
  std::lock_guard<CtxMutex> lk(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  const int64_t n2 = bytes / 16;
  DevBuf<double> a, b;
  a.alloc(ctx, (size_t)n2 * 2);
  b.alloc(ctx, (size_t)n2 * 2);
  hipStream_t s = ctx->stream;
  k_fill<double><<<nblk(n2 * 2), 256, 0, s>>>(n2 * 2, a.p, 1.0);
  const int grid = 1024; // fastest measured shape of this copy on MI355X
  k_stream_copy<<<grid, 256, 0, s>>>(n2, (const dvec2 *)a.p, (dvec2 *)b.p);
  HIPCHK(hipEventRecord(ctx->ev0, s));
  for (int r = 0; r < reps; r++) k_stream_copy<<<grid, 256, 0, s>>>(n2, (const dvec2 *)a.p, (dvec2 *)b.p);
  HIPCHK(hipEventRecord(ctx->ev1, s));
  HIPCHK(hipEventSynchronize(ctx->ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  *avg_ms = ms / reps;
  API_END
}

// the solver's tunables after construction (smoothing sweeps per side)
extern "C" int afe_amg_set_smooth(afe_amg *g, int32_t smooth) {
  if (!g || smooth < 0 || smooth > 16) return fail(AFE_E_ARG, "bad smoothing count");
  g->prm.smooth = smooth;
  return AFE_OK;
}

// AMGSolver::isSolveMostProjectExactly() (AMGSolver.h:100-101): exact = 0
// replaces the coarsest solve by smoothing-count Gauss-Seidel sweeps from zero
// (src/AMGSolver.cpp:601-605)
extern "C" int afe_amg_set_coarse_exact(afe_amg *g, int32_t exact) {
  if (!g || !g->ctx) return fail(AFE_E_ARG, "bad arguments");
  std::lock_guard<CtxMutex> lk(g->ctx->mu);
  if ((exact != 0) != g->coarse_exact) {
    drop_graphs(g);
    g->coarse_exact = exact != 0;
    set_coarse_mode(g);
  }
  return AFE_OK;
}

// live timing of the level-0 SpMV inside afe_amg_pcg: enable = 1 resets and
// starts recording, 0 stops; the average so far is returned either way
extern "C" int afe_amg_live_timing(afe_amg *g, int32_t enable, double *avg_ms, int64_t *launches) {
  API_BEGIN
  if (!g || !g->ctx) return fail(AFE_E_ARG, "bad arguments");
  std::lock_guard<CtxMutex> lk(g->ctx->mu);
  HIPCHK(hipSetDevice(g->ctx->device));
  if (avg_ms) *avg_ms = g->live_n ? g->live_ms / g->live_n : 0.0;
  if (launches) *launches = g->live_n;
  if (enable && !g->live0) {
    HIPCHK(hipEventCreate(&g->live0));
    HIPCHK(hipEventCreate(&g->live1));
  }
  if (enable) {
    g->live_ms = 0;
    g->live_n = 0;
    g->live_missed = 0;
  }
  if ((enable != 0) != g->live) {
    g->live = enable != 0;
    drop_graphs(g); // the captured iteration loses or regains its SpMV launch
  }
  API_END
}

// graphs-on timing of the two parts of every CG iteration of afe_amg_pcg
// (HIP events recorded between the captured graphs, outside them)
extern "C" int afe_amg_part_timing(afe_amg *g, int32_t enable, double *part0_ms, double *part1_ms,
                                   int64_t *iterations) {
  API_BEGIN
  if (!g || !g->ctx) return fail(AFE_E_ARG, "bad arguments");
  std::lock_guard<CtxMutex> lk(g->ctx->mu);
  HIPCHK(hipSetDevice(g->ctx->device));
  if (part0_ms) *part0_ms = g->part_n ? g->part_ms[0] / g->part_n : 0.0;
  if (part1_ms) *part1_ms = g->part_n ? g->part_ms[1] / g->part_n : 0.0;
  if (iterations) *iterations = g->part_n;
  if (enable && !g->pev[0][0])
    for (auto &set : g->pev)
      for (auto &e : set) HIPCHK(hipEventCreate(&e));
  if (enable) {
    g->part_ms[0] = g->part_ms[1] = 0;
    g->part_n = 0;
  }
  g->parts = enable != 0;
  g->part_pending = false;
  API_END
}

// levels of at most `rows` rows run in the single-workgroup tail from now on
extern "C" int afe_amg_set_tail_rows(afe_amg *g, int64_t rows) {
  if (!g || !g->ctx || rows < 0) return fail(AFE_E_ARG, "bad arguments");
  std::lock_guard<CtxMutex> lk(g->ctx->mu);
  const int L = (int)g->lv.size();
  drop_graphs(g);
  g->tail0 = L - 1;
  while (g->tail0 > 0 && g->lv[g->tail0 - 1].n <= rows) g->tail0--;
  return AFE_OK;
}

// per-level device layout summary: rows, colours, stored sliced-ELL entries
extern "C" int afe_amg_device_info(const afe_amg *g, int32_t level, int64_t *rows, int32_t *colours,
                                   int64_t *stored, int32_t *tail_level, int32_t *row_lanes) {
  if (!g || !g->ctx || level < 0 || level >= (int)g->lv.size()) return fail(AFE_E_ARG, "bad level");
  const DevLevel &d = g->lv[level];
  if (rows) *rows = d.n;
  if (colours) *colours = d.ncol;
  if (stored) *stored = (int64_t)d.cols.n;
  if (tail_level) *tail_level = g->tail0;
  if (row_lanes) *row_lanes = 1 << d.gs;
  return AFE_OK;
}

// launch tuning (process-wide): xcd = XCD-aware block order (0/1),
// spmv_blocks = grid of the grid-stride level-0 kernels (8..16384, multiple of 8)
extern "C" int afe_set_tuning(int32_t xcd, int32_t spmv_blocks, int64_t tail_rows) {
  if (xcd < 0 || xcd > 2 || spmv_blocks < 8 || spmv_blocks > kMaxGridBlocks || spmv_blocks % 8 || tail_rows < 0)
    return fail(AFE_E_ARG, "bad tuning values");
  g_tune.xcd = xcd;
  g_tune.spmv_blocks = spmv_blocks;
  g_tune.tail_rows = tail_rows;
  g_tune.epoch++;
  return AFE_OK;
}

#include "batch_impl.hpp"
#include "server_impl.hpp"
