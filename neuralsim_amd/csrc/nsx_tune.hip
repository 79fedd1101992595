// nsx_tune.hip -- launch-shape autotuner.
//
// Chooses, by measurement on the context's own network and state, the launch parameters the
// step's cost depends on beyond the network itself: the propagation split (wavefronts that
// share one neuron block's spike segment), the row-split threshold that cuts high-degree rows
// into pieces those wavefronts share, the synapse pass's workgroup size, the point models'
// neurons per update thread, the graph window (steps per captured hipGraph launch), the
// delivery (ring atomics or binned, with its tile size) and persistent stepping.  The neuron block of the spike
// segments stays 256 neurons.  Every trial times the same steps from a
// device-side snapshot of the mutable state (after its warm-up / capture the state is restored)
// and the state is restored afterwards, so tuning does not change the simulation.
#include <hip/hip_runtime.h>

#include <vector>

#include "nsx_internal.h"

namespace nsx {

struct Buf {
  void* p;
  size_t bytes;
};

static std::vector<Buf> mutable_buffers(nsx_ctx* ctx) {
  const size_t n = ctx->n, nblk = ctx->p.nblk;
  DevPtrs& d = ctx->d;
  std::vector<Buf> b = {
      {d.v, n * 4},       {d.u, n * 4},       {d.gm, n * 4},        {d.gh, n * 4},
      {d.gn, n * 4},      {d.s0, n * 4},      {d.s1, n * 4},        {d.s2, n * 4},
      {d.drv, n * 4},     {d.refr, n * 4},    {d.last, n * 4},      {d.count, n * 4},
      {d.pre, n * sizeof(TraceRec)}, {d.post, n * sizeof(PostRec)},
      {d.ring, (size_t)ctx->D * ctx->ch * n * 8},
      {d.spikes, nblk * kBlock * 4},          {d.bcnt, nblk * 4},
      {d.part, nblk * kNumPart * 8},          {d.counters, kNumCounters * 8},
      {d.step_counts, (size_t)ctx->p.step_counts_cap * 4},
      {d.tbase, 64},          {d.ev_last, kEvStripes * kEvStride * 4}};
  if (ctx->cfg.plasticity != NSX_PLAST_NONE) b.push_back({ctx->w, ctx->m * 4});
  if (ctx->sur.fb) b.push_back({ctx->sur.fb, nblk * 8});
  return b;
}

static hipError_t copy_all(nsx_ctx* ctx, const std::vector<Buf>& bufs, char* snap, bool save) {
  size_t off = 0;
  for (const Buf& b : bufs) {
    hipError_t e = save ? hipMemcpyAsync(snap + off, b.p, b.bytes, hipMemcpyDeviceToDevice,
                                         ctx->stream)
                        : hipMemcpyAsync(b.p, snap + off, b.bytes, hipMemcpyDeviceToDevice,
                                         ctx->stream);
    if (e != hipSuccess) return e;
    off += (b.bytes + 255) & ~(size_t)255;
  }
  return hipStreamSynchronize(ctx->stream);
}

static size_t snapshot_bytes(const std::vector<Buf>& bufs) {
  size_t total = 0;
  for (const Buf& b : bufs) total += (b.bytes + 255) & ~(size_t)255;
  return total;
}

}  // namespace nsx

using namespace nsx;

extern "C" {

nsx_status nsx_snapshot_save(nsx_ctx* ctx) {
  if (!ctx) return set_error(nullptr, NSX_E_INVALID, "ctx is NULL");
  if (!ctx->has_network) return set_error(ctx, NSX_E_STATE, "no network loaded");
  nsx_status s = bin_materialize(ctx);  // pending buckets join the ring the snapshot holds
  if (s) return s;
  const std::vector<Buf> bufs = mutable_buffers(ctx);
  const size_t total = snapshot_bytes(bufs);
  if (ctx->snap && ctx->snap_bytes != total) {
    dev_free(ctx, ctx->snap);
    ctx->snap = nullptr;
  }
  if (!ctx->snap) {
    if ((s = dev_alloc(ctx, &ctx->snap, total, false))) return s;
    ctx->snap_bytes = total;
  }
  NSX_HIP(ctx, copy_all(ctx, bufs, (char*)ctx->snap, true));
  ctx->snap_t = ctx->t;
  ctx->snap_tb = ctx->tbase_host;
  ctx->snap_hs = ctx->hs;
  return NSX_OK;
}

nsx_status nsx_snapshot_restore(nsx_ctx* ctx) {
  if (!ctx) return set_error(nullptr, NSX_E_INVALID, "ctx is NULL");
  if (!ctx->snap) return set_error(ctx, NSX_E_STATE, "no snapshot saved");
  const std::vector<Buf> bufs = mutable_buffers(ctx);
  if (snapshot_bytes(bufs) != ctx->snap_bytes)
    return set_error(ctx, NSX_E_STATE, "the network changed since the snapshot was saved");
  NSX_HIP(ctx, copy_all(ctx, bufs, (char*)ctx->snap, false));
  ctx->t = ctx->snap_t;
  ctx->tbase_host = ctx->snap_tb;
  ctx->hs = ctx->snap_hs;
  return bin_clear(ctx);  // events emitted after the save are not part of it
}

nsx_status nsx_autotune(nsx_ctx* ctx, uint32_t trial_steps, nsx_tune_report* rep) {
  if (!ctx) return set_error(nullptr, NSX_E_INVALID, "ctx is NULL");
  if (!ctx->has_network) return set_error(ctx, NSX_E_STATE, "no network loaded");
  if (trial_steps < 1) return set_error(ctx, NSX_E_INVALID, "trial_steps must be >= 1");
  if (ctx->cfg.model == NSX_MODEL_HH_SURROGATE && !(ctx->sur.has_params && ctx->sur.has_norm))
    return set_error(ctx, NSX_E_STATE, "surrogate parameters not set");
  nsx_status s0 = bin_materialize(ctx);  // the snapshot's ring then holds all pending input
  if (s0) return s0;
  if ((s0 = epoch_prepare(ctx))) return s0;
  const std::vector<Buf> bufs = mutable_buffers(ctx);
  const size_t total = snapshot_bytes(bufs);
  char* snap = nullptr;
  nsx_status s = dev_alloc(ctx, (void**)&snap, total, false);
  if (s) return s;
  NSX_HIP(ctx, copy_all(ctx, bufs, snap, true));
  const uint64_t t0 = ctx->t;
  const int32_t tb0 = ctx->tbase_host;
  const bool prof = ctx->profiling;
  const uint32_t split0 = ctx->p.prop_split, graph0 = ctx->graph_steps;
  const uint32_t deliv0 = ctx->delivery, te0 = ctx->bin_te_req;
  const uint32_t rs0 = ctx->p.row_split, nt0 = ctx->syn_block, width0 = ctx->update_width;
  const uint64_t cap0 = ctx->bin_cap_req;
  const bool persist0 = ctx->persistent;
  const uint32_t stats0 = ctx->p.stats_on;  // trials neither count nor fold statistics
  const bool point = ctx->cfg.model == NSX_MODEL_LIF || ctx->cfg.model == NSX_MODEL_IZHIKEVICH;
  const bool per_step = epoch_len(ctx) == 1;  // an epoch launch has its own update and block
  ctx->profiling = false;
  ctx->p.stats_on = 0;
  destroy_graph(ctx);
  nsx_tune_report r{};
  double best = 1e300;
  struct Shape {
    uint32_t split, graph, dl, te, rs, nt, width;  // dl: delivery; te: binned tile entries
    uint32_t ps;                                   // persistent stepping
  };
  Shape cur{split0, graph0 ? graph0 : 50, deliv0,
            deliv0 == NSX_DELIVERY_BINNED ? ctx->bin_te : 0u, rs0, nt0, width0, 0u};
  Shape bs = cur;
  auto trial = [&](const Shape& sh) -> nsx_status {
    nsx_status st =
        bin_configure(ctx, sh.dl, cap0, sh.dl == NSX_DELIVERY_BINNED ? sh.te : 0u);
    if (!st) st = nsx_set_prop_split(ctx, sh.split);
    if (!st) st = nsx_set_row_split(ctx, sh.rs);
    if (!st) st = nsx_set_syn_block(ctx, sh.nt);
    if (!st) st = nsx_set_update_width(ctx, sh.width);
    if (!st) st = nsx_set_graph_steps(ctx, sh.graph);
    if (!st) st = nsx_set_persistent(ctx, (int)sh.ps);
    if (st) return st;
    // the same number of steps in every trial: a multiple of every window tried (200)
    const uint64_t steps = (trial_steps + 199) / 200 * 200;
    nsx_step_stats ss{};
    if ((st = nsx_step(ctx, sh.graph ? sh.graph : 8, &ss))) return st;  // capture / warm up
    // every trial times the same steps: back to the snapshot (the captured graph stays valid:
    // the buffers keep their addresses and the step base is part of the snapshot)
    hipError_t e0 = copy_all(ctx, bufs, snap, false);
    if (e0 != hipSuccess) return hip_error(ctx, e0, "autotune restore");
    ctx->t = t0;
    ctx->tbase_host = tb0;
    if ((st = bin_clear(ctx))) return st;
    if ((st = nsx_step(ctx, steps, &ss))) return st;
    const double us = ss.wall_ms * 1000.0 / (double)steps;
    if (r.n_trials < 48) {
      const uint32_t k = r.n_trials++;
      r.trial_split[k] = sh.split;
      r.trial_graph[k] = sh.graph;
      r.trial_tile[k] = sh.dl == NSX_DELIVERY_BINNED ? sh.te : 0u;
      r.trial_delivery[k] = sh.dl;
      r.trial_row_split[k] = sh.rs;
      r.trial_syn_block[k] = sh.nt;
      r.trial_width[k] = sh.width;
      r.trial_persistent[k] = sh.ps;
      r.us_per_step[k] = us;
    }
    if (us < best) {
      best = us;
      bs = sh;
    }
    hipError_t e = copy_all(ctx, bufs, snap, false);
    if (e != hipSuccess) return hip_error(ctx, e, "autotune restore");
    ctx->t = t0;
    ctx->tbase_host = tb0;
    return bin_clear(ctx);  // the trial's events in flight are not part of the snapshot
  };
  static const uint32_t kSplits[] = {1, 2, 4, 8, 16};
  static const uint32_t kGraphs[] = {10, 25, 100, 200};
  // 1. the split at a 50-step window and the current delivery
  for (uint32_t sp : kSplits) {
    Shape sh = cur;
    sh.split = sp;
    sh.graph = 50;
    if (!s) s = trial(sh);
  }
  // 2. atomic delivery: the row-split threshold for high-degree rows, the workgroup size
  if (!s && bs.dl == NSX_DELIVERY_ATOMIC) {
    const Shape base = bs;
    for (uint32_t rs : {256u, 128u, 64u}) {
      Shape sh = base;
      sh.rs = rs;
      if (!s && rs != base.rs) s = trial(sh);
    }
    if (per_step) {
      const Shape base2 = bs;
      for (uint32_t nt : {128u, 256u, 512u}) {
        Shape sh = base2;
        sh.nt = nt;
        if (!s && nt != base2.nt) s = trial(sh);
      }
    }
  }
  // 3. the point models' neurons per update thread
  if (!s && point && per_step) {
    const Shape base = bs;
    for (uint32_t w : {1u, 4u}) {
      Shape sh = base;
      sh.width = w;
      if (!s) s = trial(sh);
    }
  }
  // 4. the graph window
  if (!s) {
    const Shape base = bs;
    for (uint32_t g : kGraphs) {
      Shape sh = base;
      sh.graph = g;
      if (!s) s = trial(sh);
    }
  }
  // 5. the other deliveries, and the split again under a push delivery that wins
  const uint32_t dl_start = cur.dl, te_start = cur.te;
  if (!s) {
    const Shape base = bs;
    static const uint32_t kDl[] = {NSX_DELIVERY_ATOMIC, NSX_DELIVERY_BINNED, NSX_DELIVERY_BINNED,
                                   NSX_DELIVERY_PULL};
    static const uint32_t kTe[] = {0u, 8192u, 16384u, 0u};
    for (uint32_t k = 0; k < 4; ++k) {
      const uint32_t dl = kDl[k], te = kTe[k];
      if (s || (dl == dl_start && te == te_start)) continue;
      if (dl == NSX_DELIVERY_BINNED && !bin_supported(ctx, te)) continue;
      if (dl == NSX_DELIVERY_PULL && !pull_supported(ctx)) continue;
      Shape sh = base;
      sh.dl = dl;
      sh.te = te;
      if (dl != NSX_DELIVERY_ATOMIC) sh.rs = 0;  // binned passes walk whole rows
      s = trial(sh);
    }
  }
  if (!s && bs.dl != NSX_DELIVERY_PULL && (bs.dl != dl_start || bs.te != te_start)) {
    const Shape base = bs;
    for (uint32_t sp : kSplits) {
      Shape sh = base;
      sh.split = sp;
      if (!s && sp != base.split) s = trial(sh);
    }
  }
  // 6. persistent stepping (ring atomics, one update launch per step), windows of the best
  // graph window and of 200 steps.  A device that cannot hold the grid at once refuses the
  // cooperative launch: that trial is dropped and the state restored.
  if (!s && bs.dl == NSX_DELIVERY_ATOMIC && per_step && !ctx->backend_mode &&
      ctx->cfg.model != NSX_MODEL_HH_SURROGATE) {
    const Shape base = bs;
    for (uint32_t g : {base.graph, 200u}) {
      Shape sh = base;
      sh.ps = 1u;
      sh.graph = g;
      if (g == 200u && base.graph == 200u) continue;
      if (trial(sh) != NSX_OK) {
        (void)hipGetLastError();
        hipError_t e = copy_all(ctx, bufs, snap, false);
        if (e != hipSuccess) {
          s = hip_error(ctx, e, "autotune restore");
          break;
        }
        ctx->t = t0;
        ctx->tbase_host = tb0;
        break;
      }
    }
  }
  dev_free(ctx, snap);
  ctx->profiling = prof;
  ctx->p.stats_on = stats0;
  destroy_graph(ctx);
  if (s) {
    (void)bin_configure(ctx, deliv0, cap0, te0);
    (void)nsx_set_prop_split(ctx, split0);
    (void)nsx_set_row_split(ctx, rs0);
    (void)nsx_set_syn_block(ctx, nt0);
    (void)nsx_set_update_width(ctx, width0);
    (void)nsx_set_graph_steps(ctx, graph0);
    (void)nsx_set_persistent(ctx, persist0 ? 1 : 0);
    return s;
  }
  if ((s = bin_configure(ctx, bs.dl, cap0, bs.dl == NSX_DELIVERY_BINNED ? bs.te : 0u)))
    return s;
  if ((s = nsx_set_prop_split(ctx, bs.split)) || (s = nsx_set_row_split(ctx, bs.rs)) ||
      (s = nsx_set_syn_block(ctx, bs.nt)) || (s = nsx_set_update_width(ctx, bs.width)) ||
      (s = nsx_set_graph_steps(ctx, bs.graph)) || (s = nsx_set_persistent(ctx, (int)bs.ps)))
    return s;
  r.prop_split = bs.split;
  r.graph_steps = bs.graph;
  r.delivery = bs.dl;
  r.tile_entries = bs.dl == NSX_DELIVERY_BINNED ? bs.te : bs.dl == NSX_DELIVERY_PULL ? kPullTE : 0u;
  r.row_split = bs.rs;
  r.syn_block = bs.nt;
  r.update_width = point ? bs.width : 0u;
  r.persistent = bs.ps;
  r.best_us_per_step = best;
  if (rep) *rep = r;
  return NSX_OK;
}

}  // extern "C"
