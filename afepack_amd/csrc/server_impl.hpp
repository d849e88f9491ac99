// server_impl.hpp -- a solve server on one device (part of capi.hip).
//
// Requests (structured meshes or the caller's own, optionally a moving-mesh
// update of a cached mesh) are set up by a pool of host threads -- mesh, FE
// space, device assembly, right-hand side, Dirichlet elimination and the lazy
// AMG hierarchy; for a cached mesh only the new coordinates, the values and
// AMGSolver::lazyReinit; for a new mesh whose connectivity matches a cached
// entry, a copy of that entry's space and hierarchy structure with the new
// values computed on the device -- and then admitted into one continuously running
// batch (afe_batch_*): requests that arrive while others are solving join at
// the next step, finished ones leave, and every step advances all of them
// with shared launches.  Meshes, matrices and hierarchies stay cached in HBM
// under a byte budget (least recently used idle entries are evicted).

namespace {

struct SrvEntry {
  int64_t key = 0;
  int32_t kind = 0, n = 0;
  double jitter = 0;
  uint64_t seed = 0;
  bool own_mesh = false;      // the caller's mesh (ev and vmark kept to compare)
  uint64_t struct_key = 0;    // connectivity key (kind, elements, marks)
  afe_space *sp = nullptr;
  afe_matrix *A = nullptr;
  afe_amg *g = nullptr;
  double *d_b = nullptr, *d_x = nullptr;
  std::vector<double> base_coords;
  std::vector<int32_t> vmark, ev;
  int64_t n_dof = 0, est_bytes = 0;
  bool busy = false;          // being set up or solved
  bool in_setup = false;      // its space / hierarchy being written
  int pins = 0;               // set-ups copying its structure right now
  uint64_t last_use = 0;
};

struct SrvRequest {
  int64_t id = 0;
  bool moving = false; // set up as a moving-mesh re-solve of a cached mesh
  bool cloned = false; // set up from a cached entry of the same connectivity
  afe_request r{};
  std::vector<double> fk, fphi, fw; // owned copies of the coefficient field
  std::vector<double> mcoords;      // owned copies of the caller's mesh
  std::vector<int32_t> mev, mvmark;
  uint64_t struct_key = 0;
  SrvEntry *tmpl = nullptr;         // entry whose structure the set-up copies
  int state = 0;                    // 0 pending, 1 setup, 2 ready, 3 solving, 4 done
  int status = AFE_OK;
  std::string err; // message of a failed set-up or solve
  SrvEntry *e = nullptr; // the cache entry while the request runs (null once done: it may be evicted)
  int64_t n_dof = 0;      // DOFs of the request's space, kept when it finishes
  int32_t slot = -1;
  std::vector<double> x; // solution (when keep_x)
  afe_solve_stats st{};
  // t_taken: a set-up thread took it; t_setup: HBM room found, set-up starts
  std::chrono::steady_clock::time_point t_submit, t_taken, t_setup, t_ready, t_admit, t_done;
  double cpu_user = 0, cpu_sys = 0, lock_wait = 0; // of the set-up thread during the set-up
};

// user and system CPU seconds of the calling thread
static void thread_cpu(double &user, double &sys) {
  struct rusage ru;
  if (getrusage(RUSAGE_THREAD, &ru) != 0) {
    user = sys = 0;
    return;
  }
  user = ru.ru_utime.tv_sec + 1e-6 * ru.ru_utime.tv_usec;
  sys = ru.ru_stime.tv_sec + 1e-6 * ru.ru_stime.tv_usec;
}

// structured-mesh sizes
static void mesh_sizes(int kind, int n, int64_t &nv, int64_t &ne, int64_t &ndof, int64_t &nnz) {
  const int64_t m = n;
  if (kind == AFE_P1_TET) {
    nv = (m + 1) * (m + 1) * (m + 1);
    ne = 6 * m * m * m;
    ndof = nv;
    nnz = 15 * m * m * m + 21 * m * m + 9 * m + 1;
  } else {
    nv = (m + 1) * (m + 1);
    ne = 2 * m * m;
    ndof = kind == AFE_P2_TRI ? (2 * m + 1) * (2 * m + 1) : nv;
    nnz = kind == AFE_P2_TRI ? 46 * m * m + 16 * m + 1 : 7 * m * m + 6 * m + 1;
  }
}

// HBM of one cached problem before it is built (reserved for admission; the
// measured size replaces it once built): pattern (int64), values, gather
// map, the hierarchy's sliced-ELL copies (FP64, FP32, int32 columns, natural
// copy), refresh maps, transfer operators and about forty vectors
static int64_t entry_bytes(int kind, int n) {
  int64_t nv, ne, ndof, nnz;
  mesh_sizes(kind, n, nv, ne, ndof, nnz);
  return 96 * nnz + 480 * ndof + 96 * ne;
}

// the same for a caller's mesh, with upper bounds from its element count:
// at most n_loc^2 stored entries per element, one DOF per vertex plus (P2)
// three edges per element
static int64_t entry_bytes_mesh(int kind, int64_t nv, int64_t ne) {
  const int nl = kind == AFE_P2_TRI ? 6 : kind == AFE_P1_TET ? 4 : 3;
  const int64_t ndof = nv + (kind == AFE_P2_TRI ? 3 * ne : 0);
  return 96 * (int64_t)nl * nl * ne + 480 * ndof + 96 * ne;
}

// connectivity keys: FNV-1a over the words that fix the numbering, pattern
// and lazy hierarchy (equal keys are confirmed by comparing the arrays)
static uint64_t fnv(uint64_t h, const void *p, size_t bytes) {
  const unsigned char *c = (const unsigned char *)p;
  for (size_t i = 0; i < bytes; i++) {
    h ^= c[i];
    h *= 1099511628211ull;
  }
  return h;
}
static uint64_t struct_key_structured(int kind, int n) {
  uint64_t h = 1469598103934665603ull;
  const int32_t w[3] = {0x5354, kind, n};
  return fnv(h, w, sizeof w);
}
static uint64_t struct_key_mesh(int kind, const std::vector<int32_t> &ev, const std::vector<int32_t> &vm) {
  uint64_t h = 1469598103934665603ull;
  const int64_t w[3] = {0x4d45, kind, (int64_t)vm.size()};
  h = fnv(h, w, sizeof w);
  h = fnv(h, ev.data(), ev.size() * sizeof(int32_t));
  return fnv(h, vm.data(), vm.size() * sizeof(int32_t));
}

} // namespace

struct afe_server {
  afe_ctx *ctx = nullptr;
  afe_batch *batch = nullptr;
  int64_t budget = 0, reserved = 0;
  int max_slots = 0;
  std::map<int64_t, std::unique_ptr<SrvEntry>> cache;
  std::map<int64_t, std::unique_ptr<SrvRequest>> reqs;
  std::deque<SrvRequest *> pending, ready;
  std::mutex m;
  std::condition_variable cv_work, cv_done;
  std::vector<std::thread> workers;
  std::thread engine;
  bool stop = false;
  // wall seconds of the set-up steps summed over requests (AFE_VERBOSE
  // prints their means at destroy): new meshes 0-7, moving meshes 8-14
  double t_phase[15] = {};
  int64_t n_phase_new = 0, n_phase_moving = 0;
  // seconds the server lock was held for evictions and for the engine's
  // batch calls (AFE_VERBOSE)
  double t_evict = 0, t_engine_locked = 0;
  int64_t next_id = 1, completed = 0, evictions = 0;
  uint64_t clock = 0;
  std::string engine_error;
  // sums over finished requests: set-up of new / moved meshes, waiting for a
  // set-up thread, waiting for a batch slot, solving; counts of new / moved
  double t_setup_new = 0, t_setup_moving = 0, t_wait_setup = 0, t_wait_slot = 0, t_solve = 0;
  int64_t n_new = 0, n_moving = 0;
  double t_setup_cloned = 0; // the part of t_setup_new spent on meshes set up by structure copies
  int64_t n_cloned = 0;
  // waiting for HBM room before a set-up; the set-up threads' user and
  // system CPU seconds and context-lock waits during set-ups
  double t_wait_room = 0, cpu_user = 0, cpu_sys = 0, lock_wait = 0;
};

static void srv_free_entry(afe_server *s, SrvEntry *e) {
  if (e->g) afe_amg_destroy(e->g);
  if (e->A) afe_matrix_destroy(e->A);
  if (e->sp) afe_space_destroy(e->sp);
  if (e->d_b) afe_dev_free(s->ctx, e->d_b);
  if (e->d_x) afe_dev_free(s->ctx, e->d_x);
  e->g = nullptr; e->A = nullptr; e->sp = nullptr; e->d_b = e->d_x = nullptr;
}

// device bytes held: the larger of the cache's accounted sizes and the
// context's live allocations, plus the reservations of in-flight set-ups
static int64_t srv_bytes(afe_server *s) {
  int64_t b = 0;
  for (auto &kv : s->cache)
    if (kv.second->g) b += kv.second->est_bytes;
  int64_t live;
  {
    std::lock_guard<CtxMutex> lc(s->ctx->mu);
    live = s->ctx->bytes;
  }
  return std::max(b, live) + s->reserved;
}

// make room for `need` bytes: evict idle entries, least recently used first
// (caller holds s->m); false when only busy entries are left
static bool srv_make_room(afe_server *s, int64_t need) {
  while (srv_bytes(s) + need > s->budget) {
    SrvEntry *victim = nullptr;
    for (auto &kv : s->cache) {
      SrvEntry *e = kv.second.get();
      if (e->busy || e->pins > 0 || !e->g) continue;
      if (!victim || e->last_use < victim->last_use) victim = e;
    }
    if (!victim) return false;
    const auto t0 = std::chrono::steady_clock::now();
    srv_free_entry(s, victim);
    s->t_evict += seconds_between(t0, std::chrono::steady_clock::now());
    s->evictions++;
    s->cache.erase(victim->key);
  }
  return true;
}

// the request's mesh: the structured generator (then, for a re-solve, the
// moving-mesh displacement of the base coordinates) or the caller's copy;
// the elements only for a new entry (want_ev)
static int srv_mesh(SrvEntry *e, const SrvRequest *q, std::vector<double> &c, std::vector<int32_t> &ev,
                    bool want_ev) {
  const afe_request &r = q->r;
  const int dim = r.kind == AFE_P1_TET ? 3 : 2;
  if (e->own_mesh) {
    c = q->mcoords;
    ev = q->mev;
    if (e->base_coords.empty()) {
      e->vmark = q->mvmark;
      e->base_coords = q->mcoords;
    }
    return AFE_OK;
  }
  int64_t nv, ne, ndof, nnz;
  mesh_sizes(r.kind, r.n, nv, ne, ndof, nnz);
  if (e->base_coords.empty() || want_ev) {
    std::vector<double> bc((size_t)nv * dim);
    std::vector<int32_t> vm(nv);
    ev.resize((size_t)ne * (dim + 1));
    const int rc = r.kind == AFE_P1_TET ? afe_mesh_cube(r.n, r.jitter, r.seed, bc.data(), ev.data(), vm.data())
                                        : afe_mesh_square(r.n, r.jitter, r.seed, bc.data(), ev.data(), vm.data());
    if (rc) return rc;
    if (e->base_coords.empty()) {
      e->base_coords = std::move(bc);
      e->vmark = std::move(vm);
    }
  }
  c = e->base_coords;
  if (r.move_amp != 0) afe_mesh_displace(dim, nv, c.data(), e->vmark.data(), r.move_amp / r.n, r.move_k);
  return AFE_OK;
}

static int srv_setup(afe_server *s, SrvRequest *q) {
  SrvEntry *e = q->e;
  const afe_request &r = q->r;
  afe_field coef = r.coef;
  coef.k = q->fk.empty() ? nullptr : q->fk.data();
  coef.phi = q->fphi.empty() ? nullptr : q->fphi.data();
  coef.w = q->fw.empty() ? nullptr : q->fw.data();
  const afe_field *cf = coef.n_modes > 0 ? &coef : nullptr;
  int rc;
  std::vector<double> c;
  std::vector<int32_t> ev;
  double ph[15] = {};
  auto tp = std::chrono::steady_clock::now();
  auto mark = [&](int i) {
    const auto now = std::chrono::steady_clock::now();
    ph[i] += std::chrono::duration<double>(now - tp).count();
    tp = now;
  };
  struct Flush { // adds this set-up's step times to the server's sums on every return
    afe_server *s; double *ph; bool moving;
    ~Flush() {
      std::lock_guard<std::mutex> lk(s->m);
      for (int i = 0; i < 15; i++) s->t_phase[i] += ph[i];
      (moving ? s->n_phase_moving : s->n_phase_new)++;
    }
  } flush{s, ph, e->g != nullptr};
  if (!e->g) {
    if ((rc = srv_mesh(e, q, c, ev, true))) return rc;
    mark(0);
    const int64_t nv = (int64_t)e->vmark.size();
    const int64_t ne = (int64_t)ev.size() / (r.kind == AFE_P1_TET ? 4 : 3);
    if (q->tmpl) // same connectivity as a cached entry: copy its numbering, pattern and gather map
      rc = afe_space_clone(q->tmpl->sp, c.data(), &e->sp);
    else
      rc = afe_space_create(s->ctx, r.kind, nv, c.data(), ne, ev.data(), e->vmark.data(), &e->sp);
    if (rc) return rc;
    mark(1);
    if ((rc = afe_matrix_create(e->sp, &e->A))) return rc;
    int32_t nl;
    int64_t nz;
    afe_space_info(e->sp, &e->n_dof, &nz, &nl);
    if ((rc = afe_dev_alloc(s->ctx, e->n_dof * 8, (void **)&e->d_b))) return rc;
    if ((rc = afe_dev_alloc(s->ctx, e->n_dof * 8, (void **)&e->d_x))) return rc;
    mark(2);
    if ((rc = afe_assemble(e->A, r.op, r.acc, cf))) return rc;
    mark(3);
    if ((rc = afe_rhs(e->sp, r.acc, r.fkind, cf, nullptr, e->d_b))) return rc;
    mark(4);
    if ((rc = afe_apply_dirichlet(e->A, nullptr, e->d_b))) return rc;
    mark(5);
    if (q->tmpl) { // its hierarchy's structure, with this matrix's Galerkin products
      if ((rc = afe_amg_clone(q->tmpl->g, e->A, &e->g))) return rc;
      afe_amg_set_smooth(e->g, 1);
    } else {
      afe_amg_params p{0.25, 50u, 0.382, 1, 1, 1};
      if ((rc = afe_amg_create(e->A, &p, &e->g))) return rc;
    }
    afe_amg_set_tail_rows(e->g, g_tune.tail_rows);
    mark(6);
    rc = afe_dev_zero(s->ctx, e->d_x, e->n_dof * 8);
    mark(7);
    return rc;
  }
  // moving-mesh re-solve: new coordinates, same connectivity and pattern;
  // values-only assembly and the lazy hierarchy refresh
  if ((rc = srv_mesh(e, q, c, ev, false))) return rc;
  mark(8);
  if ((rc = afe_space_update_coords(e->sp, c.data()))) return rc;
  mark(9);
  if ((rc = afe_assemble(e->A, r.op, r.acc, cf))) return rc;
  mark(10);
  if ((rc = afe_rhs(e->sp, r.acc, r.fkind, cf, nullptr, e->d_b))) return rc;
  mark(11);
  if ((rc = afe_apply_dirichlet(e->A, nullptr, e->d_b))) return rc;
  mark(12);
  if ((rc = afe_amg_lazy_reinit(e->g, e->A))) return rc;
  mark(13);
  rc = afe_dev_zero(s->ctx, e->d_x, e->n_dof * 8);
  mark(14);
  return rc;
}

// the cached entry holds this request's mesh (same key and connectivity,
// and for a structured mesh the same base mesh)
static bool srv_same_mesh(const SrvEntry *e, const SrvRequest *q) {
  if (!e->g || e->kind != q->r.kind || e->struct_key != q->struct_key) return false;
  if (q->r.coords) return e->own_mesh && e->ev == q->mev && e->vmark == q->mvmark;
  return !e->own_mesh && e->n == q->r.n && e->jitter == q->r.jitter && e->seed == q->r.seed;
}

// an idle-or-solving entry of the same connectivity whose structure a new
// mesh may copy (not while its own space or hierarchy is being written)
static SrvEntry *srv_template(afe_server *s, const SrvRequest *q) {
  for (auto &kv : s->cache) {
    SrvEntry *t = kv.second.get();
    if (!t->g || t->in_setup || t->kind != q->r.kind || t->struct_key != q->struct_key) continue;
    if (q->r.coords ? (!t->own_mesh || t->ev != q->mev || t->vmark != q->mvmark) : (t->own_mesh || t->n != q->r.n))
      continue;
    return t;
  }
  return nullptr;
}

// true when something other than entry e holds or reserves device memory
// that will be given back (another set-up or a solve in progress)
static bool srv_others_in_flight(afe_server *s, const SrvEntry *e) {
  if (s->reserved > 0) return true;
  for (auto &kv : s->cache)
    if (kv.second.get() != e && kv.second->busy) return true;
  return false;
}

static void srv_finish_failed(afe_server *s, SrvRequest *q, SrvEntry *e, int rc, const std::string &msg) {
  q->status = rc;
  q->err = msg;
  q->state = 4;
  q->t_done = std::chrono::steady_clock::now();
  q->n_dof = e->n_dof;
  q->e = nullptr;
  if (q->tmpl) {
    q->tmpl->pins--;
    q->tmpl = nullptr;
  }
  srv_free_entry(s, e);
  s->cache.erase(e->key);
  s->completed++;
  s->cv_done.notify_all();
  s->cv_work.notify_all();
}

static void srv_worker(afe_server *s) {
  // the pool works on many requests at once: one host thread per set-up,
  // first-fit colouring (the host is the server's bottleneck)
  set_host_threads(1);
  t_first_fit_only = true;
  std::unique_lock<std::mutex> lk(s->m);
  while (!s->stop) {
    // first pending request whose mesh is not being solved, set up or copied
    SrvRequest *q = nullptr;
    for (auto it = s->pending.begin(); it != s->pending.end(); ++it) {
      auto ce = s->cache.find((*it)->r.mesh_key);
      if (ce != s->cache.end() && (ce->second->busy || ce->second->pins > 0)) continue;
      q = *it;
      s->pending.erase(it);
      break;
    }
    if (!q) {
      s->cv_work.wait(lk);
      continue;
    }
    q->state = 1;
    q->t_taken = std::chrono::steady_clock::now();
    auto ce = s->cache.find(q->r.mesh_key);
    SrvEntry *e;
    bool fresh = false;
    if (ce != s->cache.end() && srv_same_mesh(ce->second.get(), q)) {
      e = ce->second.get();
    } else {
      if (ce != s->cache.end()) { // same key, different mesh: replace
        srv_free_entry(s, ce->second.get());
        s->cache.erase(ce);
      }
      std::unique_ptr<SrvEntry> ne(new SrvEntry);
      ne->key = q->r.mesh_key;
      ne->kind = q->r.kind;
      ne->n = q->r.n;
      ne->jitter = q->r.jitter;
      ne->seed = q->r.seed;
      ne->own_mesh = q->r.coords != nullptr;
      ne->struct_key = q->struct_key;
      if (ne->own_mesh) ne->ev = q->mev;
      ne->est_bytes = ne->own_mesh ? entry_bytes_mesh(q->r.kind, (int64_t)q->mvmark.size(), (int64_t)q->mev.size() /
                                                      (q->r.kind == AFE_P1_TET ? 4 : 3))
                                   : entry_bytes(q->r.kind, q->r.n);
      e = ne.get();
      s->cache[ne->key] = std::move(ne);
      fresh = true;
      // a cached entry of the same connectivity lends its structure
      if (SrvEntry *t = srv_template(s, q)) {
        t->pins++;
        q->tmpl = t;
        e->est_bytes = t->est_bytes; // measured on the template
      }
    }
    e->busy = true;
    e->in_setup = true;
    q->e = e;
    q->moving = !fresh;
    int64_t need = fresh ? e->est_bytes : 0;
    // HBM admission: wait until the budget has room (finished requests free
    // their entries for eviction); when nothing else holds memory that will
    // come back, waiting cannot help: drop the template (it may be evicted
    // then) or fail the request
    bool refused = false;
    while (fresh && !s->stop && !srv_make_room(s, need)) {
      if (!srv_others_in_flight(s, e)) {
        if (q->tmpl) {
          q->tmpl->pins--;
          q->tmpl = nullptr;
          continue;
        }
        refused = true;
        break;
      }
      s->cv_done.wait(lk);
    }
    if (s->stop) break;
    if (refused) {
      srv_finish_failed(s, q, e, AFE_E_BUDGET, "request does not fit the HBM budget with every idle mesh evicted");
      continue;
    }
    q->cloned = q->tmpl != nullptr;
    s->reserved += need;
    lk.unlock();
    t_dev_bytes = 0;
    q->t_setup = std::chrono::steady_clock::now();
    double u0, s0, u1, s1;
    thread_cpu(u0, s0);
    const double w0 = t_ctx_lock_wait;
    const int rc = srv_setup(s, q);
    thread_cpu(u1, s1);
    q->cpu_user = u1 - u0;
    q->cpu_sys = s1 - s0;
    q->lock_wait = t_ctx_lock_wait - w0;
    const std::string msg = rc != AFE_OK ? std::string(afe_last_error()) : std::string();
    const int64_t kept = t_dev_bytes;
    lk.lock();
    s->reserved -= need;
    e->in_setup = false;
    if (q->tmpl) {
      q->tmpl->pins--;
      q->tmpl = nullptr;
    }
    if (fresh && rc == AFE_OK) e->est_bytes = kept; // measured from now on
    if (rc != AFE_OK) {
      srv_finish_failed(s, q, e, rc, msg);
      continue;
    }
    q->state = 2;
    q->t_ready = std::chrono::steady_clock::now();
    s->ready.push_back(q);
    s->cv_work.notify_all();
  }
}

static void srv_engine(afe_server *s) {
  std::vector<SrvRequest *> solving;
  std::unique_lock<std::mutex> lk(s->m);
  while (!s->stop) {
    // admit every ready request while slots are free
    const auto te0 = std::chrono::steady_clock::now();
    while (!s->ready.empty() && (int)solving.size() < s->max_slots) {
      SrvRequest *q = s->ready.front();
      s->ready.pop_front();
      const int rc = afe_batch_add(s->batch, q->e->g, q->e->d_x, q->e->d_b, q->r.rtol, q->r.max_it, &q->slot);
      if (rc != AFE_OK) {
        q->status = rc;
        q->err = afe_last_error();
        q->state = 4;
        q->t_done = std::chrono::steady_clock::now();
        q->n_dof = q->e->n_dof;
        q->e->busy = false;
        q->e = nullptr;
        s->completed++;
        s->cv_done.notify_all();
        continue;
      }
      q->state = 3;
      q->t_admit = std::chrono::steady_clock::now();
      solving.push_back(q);
    }
    s->t_engine_locked += seconds_between(te0, std::chrono::steady_clock::now());
    if (solving.empty()) {
      s->cv_work.wait(lk);
      continue;
    }
    lk.unlock();
    int32_t na = 0;
    const int rc = afe_batch_step(s->batch, &na);
    lk.lock();
    const auto te1 = std::chrono::steady_clock::now();
    if (rc != AFE_OK) s->engine_error = afe_last_error();
    for (size_t k = 0; k < solving.size();) {
      SrvRequest *q = solving[k];
      int32_t done = 0;
      const int st = afe_batch_result(s->batch, q->slot, &done, &q->st);
      if (!done && rc == AFE_OK) {
        k++;
        continue;
      }
      q->status = rc != AFE_OK ? rc : st;
      if (q->r.keep_x) {
        q->x.resize(q->e->n_dof);
        if (afe_dev_download(s->ctx, q->x.data(), q->e->d_x, q->e->n_dof * 8) != AFE_OK && q->status == AFE_OK) {
          q->status = AFE_E_HIP;
          q->err = afe_last_error();
        }
      }
      afe_batch_remove(s->batch, q->slot);
      q->slot = -1;
      q->state = 4;
      q->t_done = std::chrono::steady_clock::now();
      (q->moving ? s->t_setup_moving : s->t_setup_new) += seconds_between(q->t_setup, q->t_ready);
      (q->moving ? s->n_moving : s->n_new)++;
      if (q->cloned) {
        s->t_setup_cloned += seconds_between(q->t_setup, q->t_ready);
        s->n_cloned++;
      }
      s->t_wait_setup += seconds_between(q->t_submit, q->t_taken);
      s->t_wait_room += seconds_between(q->t_taken, q->t_setup);
      s->cpu_user += q->cpu_user;
      s->cpu_sys += q->cpu_sys;
      s->lock_wait += q->lock_wait;
      s->t_wait_slot += seconds_between(q->t_ready, q->t_admit);
      s->t_solve += seconds_between(q->t_admit, q->t_done);
      q->n_dof = q->e->n_dof;
      q->e->busy = false;
      q->e->last_use = ++s->clock;
      q->e = nullptr;
      s->completed++;
      solving[k] = solving.back();
      solving.pop_back();
      s->cv_done.notify_all();
      s->cv_work.notify_all();
    }
    s->t_engine_locked += seconds_between(te1, std::chrono::steady_clock::now());
  }
}

extern "C" int afe_server_create(int32_t device, int64_t hbm_budget, int32_t max_slots, int32_t setup_threads,
                                 int32_t smooth, afe_server **out) {
  API_BEGIN
  if (!out || hbm_budget <= 0 || max_slots < 1 || setup_threads < 1 || setup_threads > 64 || smooth < 1)
    return fail(AFE_E_ARG, "bad server arguments");
  std::unique_ptr<afe_server> s(new afe_server);
  int rc = afe_ctx_create(device, &s->ctx);
  if (rc) return rc;
  if ((rc = afe_batch_create(s->ctx, AFE_PREC_FP32, smooth, &s->batch))) {
    afe_ctx_destroy(s->ctx);
    return rc;
  }
  s->budget = hbm_budget;
  s->max_slots = max_slots;
  afe_server *sp = s.get();
  s->engine = std::thread(srv_engine, sp);
  for (int i = 0; i < setup_threads; i++) s->workers.emplace_back(srv_worker, sp);
  *out = s.release();
  API_END
}

extern "C" int afe_server_destroy(afe_server *s) {
  if (!s) return AFE_OK;
  {
    std::lock_guard<std::mutex> lk(s->m);
    s->stop = true;
  }
  s->cv_work.notify_all();
  s->cv_done.notify_all();
  for (auto &t : s->workers) t.join();
  s->engine.join();
  if (verbose()) {
    static const char *names[15] = {"mesh", "space", "matrix+vectors", "assemble", "rhs", "dirichlet", "amg",
                                    "zero", "mesh", "coords", "assemble", "rhs", "dirichlet", "lazy_reinit",
                                    "zero"};
    fprintf(stderr, "server set-up step means (s), new meshes (%lld):", (long long)s->n_phase_new);
    for (int i = 0; i < 8; i++) fprintf(stderr, " %s %.4f", names[i], s->t_phase[i] / std::max<int64_t>(1, s->n_phase_new));
    fprintf(stderr, "\nserver set-up step means (s), moving meshes (%lld):", (long long)s->n_phase_moving);
    for (int i = 8; i < 15; i++)
      fprintf(stderr, " %s %.4f", names[i], s->t_phase[i] / std::max<int64_t>(1, s->n_phase_moving));
    fprintf(stderr, "\nserver lock held for evictions %.2f s (%lld), by the engine's batch calls %.2f s\n",
            s->t_evict, (long long)s->evictions, s->t_engine_locked);
  }
  afe_batch_destroy(s->batch);
  for (auto &kv : s->cache) srv_free_entry(s, kv.second.get());
  s->cache.clear();
  afe_ctx_destroy(s->ctx);
  delete s;
  return AFE_OK;
}

extern "C" int afe_server_submit(afe_server *s, const afe_request *r, int64_t *id) {
  API_BEGIN
  if (!s || !r || !id) return fail(AFE_E_ARG, "null argument");
  if (r->kind < AFE_P1_TRI || r->kind > AFE_P1_TET || (!r->coords && r->n < 1) || r->rtol <= 0 || r->max_it < 1 ||
      (r->op != AFE_OP_STIFF && r->op != AFE_OP_VARSTIFF) ||
      (r->fkind != AFE_F_POISSON && r->fkind != AFE_F_VARCOEF && r->fkind != AFE_F_ONE))
    return fail(AFE_E_ARG, "bad request");
  if ((r->op == AFE_OP_VARSTIFF || r->fkind == AFE_F_VARCOEF) &&
      (r->coef.n_modes < 1 || !r->coef.k || !r->coef.phi || !r->coef.w))
    return fail(AFE_E_ARG, "variable coefficient without a field");
  if (r->coords && (r->n_vert < 1 || r->n_elem < 1 || !r->elem_vert || !r->vert_mark))
    return fail(AFE_E_ARG, "a caller's mesh needs vertices, elements and vertex marks");
  const int64_t est = r->coords ? entry_bytes_mesh(r->kind, r->n_vert, r->n_elem) : entry_bytes(r->kind, r->n);
  if (est > s->budget) return fail(AFE_E_BUDGET, "request larger than the HBM budget");
  std::unique_ptr<SrvRequest> q(new SrvRequest);
  q->r = *r;
  if (r->coords) {
    const int dim = r->kind == AFE_P1_TET ? 3 : 2, nvtx = r->kind == AFE_P1_TET ? 4 : 3;
    q->mcoords.assign(r->coords, r->coords + (size_t)r->n_vert * dim);
    q->mev.assign(r->elem_vert, r->elem_vert + (size_t)r->n_elem * nvtx);
    q->mvmark.assign(r->vert_mark, r->vert_mark + (size_t)r->n_vert);
    for (int32_t v : q->mev)
      if (v < 0 || v >= r->n_vert) return fail(AFE_E_ARG, "element vertex index out of range");
    q->struct_key = struct_key_mesh(r->kind, q->mev, q->mvmark);
  } else {
    q->struct_key = struct_key_structured(r->kind, r->n);
  }
  // the copies above are the request's: the caller's arrays are not kept
  q->r.coords = nullptr;
  q->r.elem_vert = q->r.vert_mark = nullptr;
  if (r->coords) q->r.coords = q->mcoords.data();
  if (r->coef.n_modes > 0) {
    const int nm = r->coef.n_modes, dm = r->coef.dim;
    q->fk.assign(r->coef.k, r->coef.k + (size_t)nm * dm);
    q->fphi.assign(r->coef.phi, r->coef.phi + nm);
    q->fw.assign(r->coef.w, r->coef.w + nm);
  }
  q->t_submit = std::chrono::steady_clock::now();
  std::lock_guard<std::mutex> lk(s->m);
  q->id = s->next_id++;
  *id = q->id;
  s->pending.push_back(q.get());
  s->reqs[q->id] = std::move(q);
  s->cv_work.notify_all();
  API_END
}

extern "C" int afe_server_poll(afe_server *s, int64_t id, int32_t *done) {
  if (!s || !done) return fail(AFE_E_ARG, "null argument");
  std::lock_guard<std::mutex> lk(s->m);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return fail(AFE_E_ARG, "unknown request");
  *done = it->second->state == 4 ? 1 : 0;
  return AFE_OK;
}

extern "C" int afe_server_wait(afe_server *s, int64_t id, double *h_x, int64_t *n_dof, afe_solve_stats *st,
                               double *latency_s) {
  if (!s) return fail(AFE_E_ARG, "null argument");
  std::unique_lock<std::mutex> lk(s->m);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return fail(AFE_E_ARG, "unknown request");
  SrvRequest *q = it->second.get();
  while (q->state != 4 && !s->stop) s->cv_done.wait(lk);
  if (q->state != 4) return fail(AFE_E_ARG, "server stopped");
  if (n_dof) *n_dof = q->n_dof;
  if (h_x && !q->x.empty()) std::memcpy(h_x, q->x.data(), q->x.size() * sizeof(double));
  if (st) *st = q->st;
  if (latency_s) *latency_s = seconds_between(q->t_submit, q->t_done);
  if (q->status != AFE_OK) {
    if (q->status == AFE_E_NOTCONVERGED) return fail(AFE_E_NOTCONVERGED, "request did not reach its tolerance");
    const std::string why = !q->err.empty() ? q->err : s->engine_error;
    return fail(q->status, "request failed" + (why.empty() ? std::string() : ": " + why));
  }
  return AFE_OK;
}

extern "C" int afe_server_release(afe_server *s, int64_t id) {
  if (!s) return fail(AFE_E_ARG, "null argument");
  std::lock_guard<std::mutex> lk(s->m);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return fail(AFE_E_ARG, "unknown request");
  if (it->second->state != 4) return fail(AFE_E_ARG, "request still running");
  s->reqs.erase(it);
  return AFE_OK;
}

extern "C" int afe_server_info(afe_server *s, int64_t *device_bytes, int64_t *cached, int64_t *completed,
                               int64_t *evictions) {
  if (!s) return fail(AFE_E_ARG, "null argument");
  std::lock_guard<std::mutex> lk(s->m);
  if (device_bytes) {
    std::lock_guard<CtxMutex> lc(s->ctx->mu);
    *device_bytes = s->ctx->bytes; // includes the afe_dev_alloc buffers
  }
  if (cached) *cached = (int64_t)s->cache.size();
  if (completed) *completed = s->completed;
  if (evictions) *evictions = s->evictions;
  return AFE_OK;
}

// sums over finished requests (seconds): out[0] set-up of new meshes,
// out[1] set-up of moving-mesh re-solves, out[2] waiting for a set-up
// thread, out[3] waiting for a batch slot, out[4] solving; out[5], out[6]
// the counts of new and moving-mesh requests
extern "C" int afe_server_timing_n(afe_server *s, double *out, int32_t n_out) {
  if (!s || !out || n_out < 0) return fail(AFE_E_ARG, "bad arguments");
  std::lock_guard<std::mutex> lk(s->m);
  const double v[13] = {s->t_setup_new, s->t_setup_moving, s->t_wait_setup, s->t_wait_slot, s->t_solve,
                        (double)s->n_new, (double)s->n_moving, s->t_setup_cloned, (double)s->n_cloned,
                        s->t_wait_room, s->cpu_user, s->cpu_sys, s->lock_wait};
  for (int i = 0; i < n_out && i < 13; i++) out[i] = v[i];
  return AFE_OK;
}

extern "C" int afe_server_timing(afe_server *s, double *out) { return afe_server_timing_n(s, out, 7); }
