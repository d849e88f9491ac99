// batch_impl.hpp -- continuous batching of AMG-preconditioned CG solves
// (part of capi.hip: it uses the solver objects defined there).
//
// A batch holds solves on hierarchies of one context.  Every step runs one
// CG iteration (SpMV, vector updates, one symmetric V-cycle) of every active
// member with one launch per step kind and level/colour, the launches
// covering all members at once (batch_kernels.hpp).  The launch tables are
// rebuilt, and the step re-captured as a hipGraph, only when the set of
// active members changes (a member admitted or finished).  Convergence,
// restarts and the delivered-vector check follow afe_amg_pcg member by
// member on the host.

namespace {

struct BatchMember {
  afe_amg *g = nullptr;
  double *d_x = nullptr;
  const double *d_b = nullptr;
  double rtol = 0, bb = 0, thr = 0, pthr = 0, rr = 0, l1 = 0, init_l1 = 0;
  int max_it = 0, it = 0, restarts = 0, outer = 0;
  bool active = false, done = false, conv = false;
  // the next step starts the member's CG from its residual r (p = z = 0,
  // rz = 0: alpha and beta are 0, so the step sets z = M r and p = z without
  // moving x); it is not counted as an iteration.  thr is set from the first
  // step's gathered (b, b) when it is still negative.
  bool fresh = false;
  int status = AFE_OK;
  std::chrono::steady_clock::time_point t_add, t_done;
};

// one batched launch: a run of segments in the segment table
struct BLaunch {
  int seg_off = 0, nseg = 0, grid = 0;
  int lv_off = 0, lvn_off = 0; // LevelDev tables (level l, level l+1)
  int tile = 1;                // slices per wavefront (V-cycle level launches)
};

} // namespace

struct afe_batch {
  afe_ctx *ctx = nullptr;
  int prec = AFE_PREC_FP32;
  std::vector<std::unique_ptr<BatchMember>> slots;
  bool dirty = true;
  // epoch: active members and launch tables
  std::vector<int> act; // slot index of each active member
  int spmv_grid = 0, row_grid = 0;
  BLaunch spmv, rows;
  std::vector<BLaunch> down_gs, down_res, down_restrict, up_prolong, up_gs; // in issue order
  std::vector<int> down_gs_level, up_gs_level;                           // level of each gs launch
  int n_levels_multi = 0;
  DevBuf<Seg> segs;
  DevBuf<LevelDev<float>> lv32;
  DevBuf<LevelDev<double>> lv64;
  DevBuf<MemberDev<float>> md32;
  DevBuf<MemberDev<double>> md64;
  DevBuf<TailJob<float>> tj32;
  DevBuf<TailJob<double>> tj64;
  DevBuf<LevelDev<float>> tl32; // every member's level descriptors, one after another
  DevBuf<LevelDev<double>> tl64;
  DevBuf<CgScalars *> scp;
  DevBuf<int32_t> spmv_first, spmv_count, row_first, row_count;
  DevBuf<double> part;
  DevBuf<CgScalars> gath;
  CgScalars *h_gath = nullptr;
  size_t h_gath_n = 0;
  hipGraphExec_t graph = nullptr;
  hipEvent_t step_ev = nullptr; // completion of the step in flight
  int nu = 1;
  // slices per wavefront of the V-cycle level launches, by level bucket:
  // [0] level 0, [1] the coarser levels run as launches
  int tile[2] = {1, 1};
  std::recursive_mutex mu;
  // host time of the epochs (tables, capture and instantiation) and the
  // steps, printed at destruction with AFE_VERBOSE
  int n_epochs = 0;
  int64_t n_steps = 0;
  double t_epoch = 0, t_graph = 0, t_step = 0;
  ~afe_batch() {
    if (graph) (void)hipGraphExecDestroy(graph);
    if (step_ev) (void)hipEventDestroy(step_ev);
    if (h_gath) (void)hipHostFree(h_gath);
  }
};

static double seconds_between(std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
  return std::chrono::duration<double>(b - a).count();
}

// restart the member's CG from the residual in its r: the next batched step
// preconditions it (see BatchMember::fresh)
static void batch_restart(BatchMember &m, hipStream_t s) {
  afe_amg *g = m.g;
  const int64_t n = g->lv[0].n;
  k_fill<double><<<nblk(n), 256, 0, s>>>(n, g->wp.p, 0.0);
  k_fill<double><<<nblk(n), 256, 0, s>>>(n, g->wz.p, 0.0);
  k_fill<double><<<1, 64, 0, s>>>(1, &g->sc.p->rz, 0.0);
  m.fresh = true;
}

// members whose pair reached its target (or the iteration limit): check each
// pair's true residual, restart, or deliver round(x_hi + x_lo).  The members
// checked in one step share each host round trip (their kernels are issued
// together, then one synchronisation and the reads of their scalars).
static void batch_check_all(afe_batch *b, const std::vector<BatchMember *> &list) {
  if (list.empty()) return;
  hipStream_t s = b->ctx->stream;
  auto fetch = [&](const std::vector<BatchMember *> &ms) {
    for (BatchMember *m : ms)
      HIPCHK(hipMemcpyAsync(m->g->h_sc, m->g->sc.p, sizeof(CgScalars), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  };
  for (BatchMember *m : list) residual_norms(m->g, true, m->g->wb.p, m->g->wx.p, m->g->wr.p, m->g->wxl.p, s);
  fetch(list);
  std::vector<BatchMember *> round, deliver;
  for (BatchMember *m : list) {
    m->rr = m->g->h_sc->rr;
    const bool pair_conv = m->rr <= m->pthr;
    if (!pair_conv && ++m->restarts <= 20 && m->it < m->max_it)
      batch_restart(*m, s);
    else
      round.push_back(m);
  }
  while (!round.empty()) {
    for (BatchMember *m : round) {
      afe_amg *g = m->g;
      const int64_t n = g->lv[0].n;
      k_dd_round<<<nblk(n), 256, 0, s>>>(n, g->wx.p, g->wxl.p);
      residual_norms(g, true, g->wb.p, g->wx.p, g->wr.p, nullptr, s);
    }
    fetch(round);
    std::vector<BatchMember *> again;
    for (BatchMember *m : round) {
      m->rr = m->g->h_sc->rr;
      m->l1 = m->g->h_sc->l1;
      m->conv = m->rr <= m->thr;
      if (m->conv || m->it >= m->max_it || m->restarts > 20 || m->outer >= 3) {
        deliver.push_back(m);
        continue;
      }
      // rounding floor reached: tighten the pair and continue iterating
      m->outer++;
      m->pthr *= 0.25;
      if (m->rr > m->pthr)
        batch_restart(*m, s);
      else
        again.push_back(m);
    }
    round.swap(again);
  }
  if (deliver.empty()) return;
  for (BatchMember *m : deliver)
    HIPCHK(hipMemcpyAsync(m->d_x, m->g->wx.p, m->g->lv[0].n * sizeof(double), hipMemcpyDeviceToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
  const auto now = std::chrono::steady_clock::now();
  for (BatchMember *m : deliver) {
    m->active = false;
    m->done = true;
    m->status = m->conv ? AFE_OK : AFE_E_NOTCONVERGED;
    m->t_done = now;
  }
  b->dirty = true;
}

template <class T> static void member_tables(afe_amg *g, MemberDev<T> &md, std::vector<LevelDev<T>> **hl);

template <> void member_tables<float>(afe_amg *g, MemberDev<float> &md, std::vector<LevelDev<float>> **hl) {
  md.b0 = g->lv[0].b32.p;
  md.x0 = g->lv[0].x32.p;
  *hl = &g->h_ld32;
}
template <> void member_tables<double>(afe_amg *g, MemberDev<double> &md, std::vector<LevelDev<double>> **hl) {
  md.b0 = g->lv[0].b64.p;
  md.x0 = g->lv[0].x64.p;
  *hl = &g->h_ld64;
}

template <class T> static DevBuf<LevelDev<T>> &lv_buf(afe_batch *b);
template <> DevBuf<LevelDev<float>> &lv_buf<float>(afe_batch *b) { return b->lv32; }
template <> DevBuf<LevelDev<double>> &lv_buf<double>(afe_batch *b) { return b->lv64; }
template <class T> static DevBuf<MemberDev<T>> &md_buf(afe_batch *b);
template <> DevBuf<MemberDev<float>> &md_buf<float>(afe_batch *b) { return b->md32; }
template <> DevBuf<MemberDev<double>> &md_buf<double>(afe_batch *b) { return b->md64; }
template <class T> static DevBuf<TailJob<T>> &tj_buf(afe_batch *b);
template <> DevBuf<TailJob<float>> &tj_buf<float>(afe_batch *b) { return b->tj32; }
template <> DevBuf<TailJob<double>> &tj_buf<double>(afe_batch *b) { return b->tj64; }
template <class T> static DevBuf<LevelDev<T>> &tl_buf(afe_batch *b);
template <> DevBuf<LevelDev<float>> &tl_buf<float>(afe_batch *b) { return b->tl32; }
template <> DevBuf<LevelDev<double>> &tl_buf<double>(afe_batch *b) { return b->tl64; }

// launch tables of the active members
template <class T> static void batch_build_epoch(afe_batch *b) {
  afe_ctx *ctx = b->ctx;
  b->act.clear();
  for (size_t k = 0; k < b->slots.size(); k++)
    if (b->slots[k] && b->slots[k]->active) b->act.push_back((int)k);
  const int K = (int)b->act.size();
  if (b->graph) {
    (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
  b->dirty = false;
  if (K == 0) return;
  std::vector<Seg> segs;
  std::vector<LevelDev<T>> lvs;
  std::vector<MemberDev<T>> mds(K);
  std::vector<TailJob<T>> tjs(K);
  std::vector<LevelDev<T>> tls;
  std::vector<CgScalars *> scs(K);
  std::vector<int32_t> sf(K), sc(K), rf(K), rcnt(K);
  std::vector<std::vector<LevelDev<T>> *> hl(K);
  int Lmulti = 0;
  for (int k = 0; k < K; k++) {
    afe_amg *g = b->slots[b->act[k]]->g;
    MemberDev<T> &m = mds[k];
    m.n = g->lv[0].n;
    m.nat_si = g->nat_si.p;
    m.nat_cols = g->nat_cols.p;
    m.nat_c16 = g->nat_c16.p;
    m.nat_cbase = g->nat_cbase.p;
    m.nat_rowmap = g->nat_rowmap.p;
    m.iperm0 = g->iperm0.p;
    m.nat_v = g->nat_v64.p;
    m.nat_diag = g->nat_diag.p;
    m.x = g->wx.p; m.xl = g->wxl.p; m.r = g->wr.p; m.p = g->wp.p; m.q = g->wq.p; m.z = g->wz.p;
    m.sc = g->sc.p;
    member_tables<T>(g, m, &hl[k]);
    scs[k] = g->sc.p;
    const int L = (int)g->lv.size();
    tjs[k].loff = (int)tls.size();
    tls.insert(tls.end(), hl[k]->begin(), hl[k]->end());
    tjs[k].minv = g->minv.p;
    tjs[k].l0 = g->tail0;
    tjs[k].lc = L - 1;
    tjs[k].coarse_mode = g->coarse_mode;
    Lmulti = std::max(Lmulti, g->tail0);
  }
  b->n_levels_multi = Lmulti;
  // a launch over members: each member contributes slices [s0, s1) in
  // blocks of `per` units
  auto add_launch = [&](BLaunch &bl, auto range, int per) {
    bl.seg_off = (int)segs.size();
    int grid = 0;
    for (int k = 0; k < K; k++) {
      int s0, s1;
      if (!range(k, s0, s1) || s1 <= s0) continue;
      segs.push_back({grid, k, s0, s1});
      grid += (s1 - s0 + per - 1) / per;
    }
    bl.nseg = (int)segs.size() - bl.seg_off;
    bl.grid = grid;
  };
  // CG launches
  {
    int grid = 0;
    b->spmv.seg_off = (int)segs.size();
    for (int k = 0; k < K; k++) {
      afe_amg *g = b->slots[b->act[k]]->g;
      sf[k] = grid;
      sc[k] = (g->nat_ns + WPB - 1) / WPB;
      segs.push_back({grid, k, 0, g->nat_ns});
      grid += sc[k];
    }
    b->spmv.nseg = K;
    b->spmv.grid = grid;
    grid = 0;
    b->rows.seg_off = (int)segs.size();
    for (int k = 0; k < K; k++) {
      const int64_t n = mds[k].n;
      rf[k] = grid;
      rcnt[k] = (int)((n + 255) / 256);
      segs.push_back({grid, k, 0, (int32_t)n});
      grid += rcnt[k];
    }
    b->rows.nseg = K;
    b->rows.grid = grid;
  }
  // V-cycle launches of the multi-launch levels
  b->down_gs.clear(); b->down_res.clear(); b->down_restrict.clear();
  b->up_prolong.clear(); b->up_gs.clear(); b->down_gs_level.clear(); b->up_gs_level.clear();
  std::vector<int> lv_off(Lmulti + 1, 0);
  for (int l = 0; l <= Lmulti; l++) {
    lv_off[l] = (int)lvs.size();
    for (int k = 0; k < K; k++) {
      const std::vector<LevelDev<T>> &H = *hl[k];
      lvs.push_back(l < (int)H.size() ? H[l] : LevelDev<T>{});
    }
  }
  for (int l = 0; l < Lmulti; l++) {
    const int tile = b->tile[l == 0 ? 0 : 1];
    int maxcol = 0;
    for (int k = 0; k < K; k++) {
      afe_amg *g = b->slots[b->act[k]]->g;
      if (l < g->tail0) maxcol = std::max(maxcol, g->lv[l].ncol);
    }
    for (int c = 0; c < maxcol; c++) {
      BLaunch bl;
      add_launch(bl, [&](int k, int &s0, int &s1) {
        afe_amg *g = b->slots[b->act[k]]->g;
        if (l >= g->tail0 || c >= g->lv[l].ncol) return false;
        s0 = g->lv[l].colour_slice[c];
        s1 = g->lv[l].colour_slice[c + 1];
        return true;
      }, WPB * tile);
      bl.lv_off = lv_off[l];
      bl.tile = tile;
      b->down_gs.push_back(bl);
      b->down_gs_level.push_back(l);
      BLaunch bu;
      add_launch(bu, [&](int k, int &s0, int &s1) {
        afe_amg *g = b->slots[b->act[k]]->g;
        if (l >= g->tail0 || c >= g->lv[l].ncol) return false;
        const int cr = g->lv[l].ncol - 1 - c; // reverse colour order after the coarse correction
        s0 = g->lv[l].colour_slice[cr];
        s1 = g->lv[l].colour_slice[cr + 1];
        return true;
      }, WPB * tile);
      bu.lv_off = lv_off[l];
      bu.tile = tile;
      b->up_gs.push_back(bu);
      b->up_gs_level.push_back(l);
    }
    BLaunch br, bp, bt;
    add_launch(br, [&](int k, int &s0, int &s1) {
      afe_amg *g = b->slots[b->act[k]]->g;
      if (l >= g->tail0) return false;
      s0 = 0; s1 = g->lv[l].ns;
      return true;
    }, WPB * tile);
    br.lv_off = lv_off[l];
    br.tile = tile;
    add_launch(bp, [&](int k, int &s0, int &s1) {
      afe_amg *g = b->slots[b->act[k]]->g;
      if (l >= g->tail0) return false;
      s0 = 0; s1 = g->lv[l].p_ns;
      return true;
    }, WPB * tile);
    bp.lv_off = lv_off[l];
    bp.tile = tile;
    bp.lvn_off = lv_off[l + 1];
    add_launch(bt, [&](int k, int &s0, int &s1) {
      afe_amg *g = b->slots[b->act[k]]->g;
      if (l >= g->tail0) return false;
      s0 = 0; s1 = g->lv[l].t_ns;
      return true;
    }, WPB * tile);
    bt.lv_off = lv_off[l];
    bt.tile = tile;
    bt.lvn_off = lv_off[l + 1];
    b->down_res.push_back(br);
    b->down_restrict.push_back(bp);
    b->up_prolong.push_back(bt);
  }
  b->segs.upload(ctx, segs);
  lv_buf<T>(b).upload(ctx, lvs);
  md_buf<T>(b).upload(ctx, mds);
  tj_buf<T>(b).upload(ctx, tjs);
  tl_buf<T>(b).upload(ctx, tls);
  b->scp.upload(ctx, scs);
  b->spmv_first.upload(ctx, sf);
  b->spmv_count.upload(ctx, sc);
  b->row_first.upload(ctx, rf);
  b->row_count.upload(ctx, rcnt);
  const int pg = std::max(b->spmv.grid, 2 * b->rows.grid);
  b->part.alloc(ctx, (size_t)std::max(pg, 1));
  b->gath.alloc(ctx, (size_t)K);
  if (b->h_gath_n < (size_t)K) {
    if (b->h_gath) HIPCHK(hipHostFree(b->h_gath));
    b->h_gath = nullptr;
    HIPCHK(hipHostMalloc((void **)&b->h_gath, sizeof(CgScalars) * K));
    b->h_gath_n = K;
  }
  HIPCHK(hipStreamSynchronize(ctx->stream));
}

// one CG iteration of every active member (the batched form of
// issue_cg_spmv, issue_cg_update and issue_cg_precondition), then the
// scalars of all members gathered
template <class T> static void batch_issue_step(afe_batch *b, hipStream_t s) {
  const int K = (int)b->act.size();
  const Seg *S = b->segs.p;
  const MemberDev<T> *md = md_buf<T>(b).p;
  const LevelDev<T> *LV = lv_buf<T>(b).p;
  double *part = b->part.p;
  // q = A p, (p, q)
  kb_spmv_dot<T><<<b->spmv.grid, BLK, 0, s>>>(S + b->spmv.seg_off, b->spmv.nseg, md, part);
  kb_finalize<<<K, 256, 0, s>>>(b->spmv_first.p, b->spmv_count.p, 1, b->spmv.grid, part, b->scp.p, 1);
  // x, r update; (T) r into the V-cycle's level-0 b
  kb_cg_update<T><<<b->rows.grid, 256, 0, s>>>(S + b->rows.seg_off, b->rows.nseg, md, part);
  kb_finalize<<<K, 256, 0, s>>>(b->row_first.p, b->row_count.p, 1, b->rows.grid, part, b->scp.p, 4);
  // symmetric V(nu, nu) cycle
  const int L = b->n_levels_multi;
  size_t gi = 0;
  for (int l = 0; l < L; l++) {
    size_t g0 = gi;
    for (int it = 0; it < b->nu; it++) {
      for (gi = g0; gi < b->down_gs.size() && b->down_gs_level[gi] == l; gi++) {
        // a level below 0 starts from the first colour step that kb_restrict took
        if (l > 0 && it == 0 && gi == g0) continue;
        const BLaunch &bl = b->down_gs[gi];
        if (bl.grid) kb_gs<T><<<bl.grid, BLK, 0, s>>>(S + bl.seg_off, bl.nseg, LV + bl.lv_off, bl.tile);
      }
    }
    const BLaunch &br = b->down_res[l];
    if (br.grid) kb_residual<T><<<br.grid, BLK, 0, s>>>(S + br.seg_off, br.nseg, LV + br.lv_off, br.tile);
    const BLaunch &bp = b->down_restrict[l];
    if (bp.grid) kb_restrict<T><<<bp.grid, BLK, 0, s>>>(S + bp.seg_off, bp.nseg, LV + bp.lv_off, LV + bp.lvn_off,
                                                    bp.tile);
  }
  const int lds_cap = (int)b->ctx->lds_bytes;
  kb_tail<T><<<K, TAIL_THREADS, (size_t)lds_cap, s>>>(tj_buf<T>(b).p, tl_buf<T>(b).p, b->nu, lds_cap);
  for (int l = L - 1; l >= 0; l--) {
    const BLaunch &bt = b->up_prolong[l];
    if (bt.grid) kb_prolong<T><<<bt.grid, BLK, 0, s>>>(S + bt.seg_off, bt.nseg, LV + bt.lv_off, LV + bt.lvn_off,
                                                   bt.tile);
    size_t first = 0;
    while (first < b->up_gs.size() && b->up_gs_level[first] != l) first++;
    for (int it = 0; it < b->nu; it++)
      for (size_t j = first; j < b->up_gs.size() && b->up_gs_level[j] == l; j++) {
        const BLaunch &bl = b->up_gs[j];
        if (bl.grid) kb_gs<T><<<bl.grid, BLK, 0, s>>>(S + bl.seg_off, bl.nseg, LV + bl.lv_off, bl.tile);
      }
  }
  // z, (r, z) and the flexible beta, new direction
  kb_take_z<T><<<b->rows.grid, 256, 0, s>>>(S + b->rows.seg_off, b->rows.nseg, md, b->rows.grid, part);
  kb_finalize<<<K, 256, 0, s>>>(b->row_first.p, b->row_count.p, 2, b->rows.grid, part, b->scp.p, 2, 1);
  kb_p_update<T><<<b->rows.grid, 256, 0, s>>>(S + b->rows.seg_off, b->rows.nseg, md);
  kb_gather_scalars<<<(K + 63) / 64, 64, 0, s>>>(b->scp.p, K, b->gath.p);
}

template <class T>
static void batch_step_impl(afe_batch *b, int *n_active, std::unique_lock<CtxMutex> &lk) {
  afe_ctx *ctx = b->ctx;
  hipStream_t s = ctx->stream;
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  if (b->dirty) {
    batch_build_epoch<T>(b);
    b->n_epochs++;
  }
  const int K = (int)b->act.size();
  if (K == 0) {
    *n_active = 0;
    return;
  }
  const auto t1 = clk::now();
  b->t_epoch += std::chrono::duration<double>(t1 - t0).count();
  if (use_graphs()) {
    if (!b->graph) {
      b->graph = capture_graph(s, [&]() { batch_issue_step<T>(b, s); });
      b->t_graph += std::chrono::duration<double>(clk::now() - t1).count();
    }
    HIPCHK(hipGraphLaunch(b->graph, s));
  } else {
    batch_issue_step<T>(b, s);
  }
  HIPCHK(hipMemcpyAsync(b->h_gath, b->gath.p, sizeof(CgScalars) * K, hipMemcpyDeviceToHost, s));
  HIPCHK(hipEventRecord(b->step_ev, s));
  // the context is free while the step runs: other callers may queue their
  // work behind it on the stream (the capture above is complete; the batch
  // itself stays locked by b->mu)
  lk.unlock();
  const hipError_t we = hipEventSynchronize(b->step_ev);
  lk.lock();
  HIPCHK(we);
  b->n_steps++;
  b->t_step += std::chrono::duration<double>(clk::now() - t1).count();
  const std::vector<int> act = b->act;
  std::vector<BatchMember *> check;
  for (int k = 0; k < K; k++) {
    BatchMember &m = *b->slots[act[k]];
    if (m.fresh) {
      m.fresh = false;
      if (m.thr < 0) {
        m.bb = b->h_gath[k].bb;
        m.init_l1 = m.l1 = b->h_gath[k].l1;
        m.thr = m.rtol * m.rtol * m.bb;
        m.pthr = 0.25 * m.thr;
      }
    } else {
      m.it++;
    }
    m.rr = b->h_gath[k].rr;
    if (!(m.rr == m.rr)) {
      m.active = false;
      m.done = true;
      m.status = AFE_E_NOTCONVERGED;
      m.t_done = std::chrono::steady_clock::now();
      b->dirty = true;
      continue;
    }
    if (m.rr <= m.pthr || m.it >= m.max_it) check.push_back(&m);
  }
  batch_check_all(b, check);
  int na = 0;
  for (auto &p : b->slots)
    if (p && p->active) na++;
  *n_active = na;
}

extern "C" int afe_batch_create(afe_ctx *ctx, int32_t prec, int32_t smooth, afe_batch **out) {
  API_BEGIN
  if (!ctx || !out || (prec != AFE_PREC_FP32 && prec != AFE_PREC_FP64) || smooth < 1 || smooth > 16)
    return fail(AFE_E_ARG, "bad batch arguments");
  std::unique_ptr<afe_batch> b(new afe_batch);
  b->ctx = ctx;
  b->prec = prec;
  b->nu = smooth;
  {
    std::lock_guard<CtxMutex> lk(ctx->mu);
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipEventCreateWithFlags(&b->step_ev, hipEventDisableTiming));
  }
  *out = b.release();
  API_END
}

// slices per wavefront of the batched V-cycle level launches: tile_level0 on
// level 0, tile_coarse on the coarser levels run as launches (1, 2, 4 or 8);
// the launch tables are rebuilt at the next step.  Results do not depend on it.
extern "C" int afe_batch_set_tiles(afe_batch *b, int32_t tile_level0, int32_t tile_coarse) {
  auto ok = [](int t) { return t == 1 || t == 2 || t == 4 || t == 8; };
  if (!b || !ok(tile_level0) || !ok(tile_coarse)) return fail(AFE_E_ARG, "tiles must be 1, 2, 4 or 8");
  std::lock_guard<std::recursive_mutex> lk(b->mu);
  if (b->tile[0] != tile_level0 || b->tile[1] != tile_coarse) {
    b->tile[0] = tile_level0;
    b->tile[1] = tile_coarse;
    b->dirty = true;
  }
  return AFE_OK;
}

extern "C" int afe_batch_destroy(afe_batch *b) {
  if (!b) return AFE_OK;
  if (verbose())
    fprintf(stderr, "batch: %lld steps, %d epochs; host time: epochs %.1f ms, graphs %.1f ms, steps %.1f ms\n",
            (long long)b->n_steps, b->n_epochs, 1e3 * b->t_epoch, 1e3 * b->t_graph, 1e3 * b->t_step);
  {
    std::lock_guard<std::recursive_mutex> bl(b->mu);
    std::lock_guard<CtxMutex> lk(b->ctx->mu);
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
  }
  delete b;
  return AFE_OK;
}

extern "C" int afe_batch_add(afe_batch *b, afe_amg *g, double *d_x, const double *d_b, double rtol,
                             int32_t max_it, int32_t *slot) {
  API_BEGIN
  if (!b || !g || !d_x || !d_b || !slot || rtol <= 0 || max_it < 1) return fail(AFE_E_ARG, "bad arguments");
  // a hierarchy of another context on the same device joins too: the batch
  // works on its own stream, and the caller has finished with the
  // hierarchy's stream (every call returns with it synchronised)
  if (!g->ctx || g->ctx->device != b->ctx->device) return fail(AFE_E_ARG, "the hierarchy belongs to another device");
  if (b->prec == AFE_PREC_FP32 && !g->has32) return fail(AFE_E_ARG, "hierarchy built without FP32 copies");
  std::lock_guard<std::recursive_mutex> bl(b->mu);
  std::lock_guard<CtxMutex> lk(b->ctx->mu);
  for (auto &p : b->slots)
    if (p && p->active && p->g == g) return fail(AFE_E_ARG, "hierarchy already has an active solve in this batch");
  HIPCHK(hipSetDevice(b->ctx->device));
  hipStream_t s = b->ctx->stream;
  if (g->prm.smooth != b->nu) g->prm.smooth = b->nu;
  std::unique_ptr<BatchMember> m(new BatchMember);
  m->g = g;
  m->d_x = d_x;
  m->d_b = d_b;
  m->rtol = rtol;
  m->max_it = max_it;
  m->t_add = std::chrono::steady_clock::now();
  const int64_t n = g->lv[0].n;
  HIPCHK(hipMemcpyAsync(g->wb.p, d_b, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(g->wx.p, d_x, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  k_fill<double><<<nblk(n), 256, 0, s>>>(n, g->wxl.p, 0.0);
  // r = b - A x and (r, r), (b, b), |r|_1 into the member's scalars; the
  // first batched step preconditions r (no V-cycle of its own here, and the
  // thresholds come from that step's gathered scalars)
  residual_norms(g, true, g->wb.p, g->wx.p, g->wr.p, g->wxl.p, s);
  m->thr = m->pthr = -1.0;
  m->active = true;
  int idx = -1;
  for (size_t k = 0; k < b->slots.size(); k++)
    if (!b->slots[k]) {
      idx = (int)k;
      break;
    }
  if (idx < 0) {
    idx = (int)b->slots.size();
    b->slots.emplace_back();
  }
  b->slots[idx] = std::move(m);
  BatchMember &mm = *b->slots[idx];
  b->dirty = true;
  batch_restart(mm, s);
  HIPCHK(hipStreamSynchronize(s));
  *slot = idx;
  API_END
}

extern "C" int afe_batch_step(afe_batch *b, int32_t *n_active) {
  API_BEGIN
  if (!b) return fail(AFE_E_ARG, "null batch");
  std::lock_guard<std::recursive_mutex> bl(b->mu);
  std::unique_lock<CtxMutex> lk(b->ctx->mu);
  HIPCHK(hipSetDevice(b->ctx->device));
  int na = 0;
  if (b->prec == AFE_PREC_FP32) batch_step_impl<float>(b, &na, lk);
  else batch_step_impl<double>(b, &na, lk);
  if (n_active) *n_active = na;
  API_END
}

extern "C" int afe_batch_run(afe_batch *b) {
  int na = 1;
  while (na > 0) {
    const int rc = afe_batch_step(b, &na);
    if (rc != AFE_OK) return rc;
  }
  return AFE_OK;
}

extern "C" int afe_batch_result(afe_batch *b, int32_t slot, int32_t *done, afe_solve_stats *st) {
  if (!b) return fail(AFE_E_ARG, "null batch");
  std::lock_guard<std::recursive_mutex> lk(b->mu);
  if (slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot]) return fail(AFE_E_ARG, "bad slot");
  const BatchMember &m = *b->slots[slot];
  if (done) *done = m.done ? 1 : 0;
  if (st) {
    st->iterations = m.it;
    st->converged = m.conv ? 1 : 0;
    st->rel_res_l2 = m.bb > 0 ? std::sqrt(m.rr / m.bb) : 0.0;
    st->res_l1 = m.l1;
    st->rel_res_l1 = m.init_l1 > 0 ? m.l1 / m.init_l1 : 0.0;
    st->seconds = m.done ? seconds_between(m.t_add, m.t_done)
                         : seconds_between(m.t_add, std::chrono::steady_clock::now());
  }
  return m.done ? m.status : AFE_OK;
}

// forget a finished (or abandon an active) member; its slot is reused
extern "C" int afe_batch_remove(afe_batch *b, int32_t slot) {
  if (!b) return fail(AFE_E_ARG, "null batch");
  std::lock_guard<std::recursive_mutex> lk(b->mu);
  if (slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot]) return fail(AFE_E_ARG, "bad slot");
  if (b->slots[slot]->active) b->dirty = true;
  b->slots[slot].reset();
  return AFE_OK;
}
