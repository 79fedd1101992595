// nsx_tune.hip -- launch-shape autotuner.
//
// Chooses, by measurement on the context's own network and state, the launch parameters the
// step's cost depends on beyond the network itself: the propagation split (wavefronts that
// share one neuron block's spike segment -- the row-split of busy blocks with high-degree
// rows), the graph window (steps per captured hipGraph launch) and the delivery (ring atomics
// or binned, with its tile size).  The neuron block size is
// fixed at 256 by the spike-segment layout.  Every trial times the same steps from a
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
  const std::vector<Buf> bufs = mutable_buffers(ctx);
  size_t total = 0;
  for (const Buf& b : bufs) total += (b.bytes + 255) & ~(size_t)255;
  char* snap = nullptr;
  nsx_status s = dev_alloc(ctx, (void**)&snap, total, false);
  if (s) return s;
  NSX_HIP(ctx, copy_all(ctx, bufs, snap, true));
  const uint64_t t0 = ctx->t;
  const int32_t tb0 = ctx->tbase_host;
  const bool prof = ctx->profiling;
  const uint32_t split0 = ctx->p.prop_split, graph0 = ctx->graph_steps;
  const uint32_t deliv0 = ctx->delivery, te0 = ctx->bin_te_req;
  const uint64_t cap0 = ctx->bin_cap_req;
  const uint32_t stats0 = ctx->p.stats_on;  // trials neither count nor fold statistics
  ctx->profiling = false;
  ctx->p.stats_on = 0;
  destroy_graph(ctx);
  nsx_tune_report r{};
  double best = 1e300;
  uint32_t best_split = split0, best_graph = graph0 ? graph0 : 50;
  uint32_t best_te = deliv0 == NSX_DELIVERY_BINNED ? ctx->bin_te : 0u;  // 0: atomic
  // te = 0: atomic delivery, else binned with that tile size
  auto trial = [&](uint32_t split, uint32_t graph, uint32_t te) -> nsx_status {
    nsx_status st = bin_configure(ctx, te ? NSX_DELIVERY_BINNED : NSX_DELIVERY_ATOMIC, cap0, te);
    if (!st) st = nsx_set_prop_split(ctx, split);
    if (!st) st = nsx_set_graph_steps(ctx, graph);
    if (st) return st;
    // the same number of steps in every trial: a multiple of every window tried (200)
    const uint64_t steps = (trial_steps + 199) / 200 * 200;
    nsx_step_stats ss{};
    if ((st = nsx_step(ctx, graph ? graph : 8, &ss))) return st;  // capture / warm up
    // every trial times the same steps: back to the snapshot (the captured graph stays valid:
    // the buffers keep their addresses and the step base is part of the snapshot)
    hipError_t e0 = copy_all(ctx, bufs, snap, false);
    if (e0 != hipSuccess) return hip_error(ctx, e0, "autotune restore");
    ctx->t = t0;
    ctx->tbase_host = tb0;
    if ((st = bin_clear(ctx))) return st;
    if ((st = nsx_step(ctx, steps, &ss))) return st;
    const double us = ss.wall_ms * 1000.0 / (double)steps;
    if (r.n_trials < 32) {
      r.trial_split[r.n_trials] = split;
      r.trial_graph[r.n_trials] = graph;
      r.trial_tile[r.n_trials] = te;
      r.us_per_step[r.n_trials] = us;
      ++r.n_trials;
    }
    if (us < best) {
      best = us;
      best_split = split;
      best_graph = graph;
      best_te = te;
    }
    hipError_t e = copy_all(ctx, bufs, snap, false);
    if (e != hipSuccess) return hip_error(ctx, e, "autotune restore");
    ctx->t = t0;
    ctx->tbase_host = tb0;
    return bin_clear(ctx);  // the trial's events in flight are not part of the snapshot
  };
  // coordinate search: the split at a 50-step window, then the window at the best split
  static const uint32_t kSplits[] = {1, 2, 4, 8, 16};
  static const uint32_t kGraphs[] = {10, 25, 100, 200};
  const uint32_t te_start = best_te;
  for (uint32_t sp : kSplits)
    if ((s = trial(sp, 50, te_start))) break;
  const uint32_t sp_best = best_split;
  if (!s)
    for (uint32_t g : kGraphs)
      if ((s = trial(sp_best, g, te_start))) break;
  // the other delivery: atomic, and binned where (D - 1) * tiles fits
  const uint32_t sp2 = best_split, g2 = best_graph;
  for (uint32_t te : {0u, 8192u, 16384u}) {
    if (s || te == te_start) continue;
    if (te && !bin_supported(ctx, te)) continue;
    s = trial(sp2, g2, te);
  }
  dev_free(ctx, snap);
  ctx->profiling = prof;
  ctx->p.stats_on = stats0;
  destroy_graph(ctx);
  if (s) {
    (void)bin_configure(ctx, deliv0, cap0, te0);
    (void)nsx_set_prop_split(ctx, split0);
    (void)nsx_set_graph_steps(ctx, graph0);
    return s;
  }
  if ((s = bin_configure(ctx, best_te ? NSX_DELIVERY_BINNED : NSX_DELIVERY_ATOMIC, cap0,
                         best_te)))
    return s;
  if ((s = nsx_set_prop_split(ctx, best_split))) return s;
  if ((s = nsx_set_graph_steps(ctx, best_graph))) return s;
  r.prop_split = best_split;
  r.graph_steps = best_graph;
  r.delivery = best_te ? NSX_DELIVERY_BINNED : NSX_DELIVERY_ATOMIC;
  r.tile_entries = best_te;
  r.best_us_per_step = best;
  if (rep) *rep = r;
  return NSX_OK;
}

}  // extern "C"
