// nsx_api.hip -- C ABI of the simulation core: context, configuration, state, stepping.
// The device-resident counterpart of SimulationEngine (src/simulation/engine.rs:43-823) with
// GpuBackend's ownership/error conventions (src/simulation/gpu_backend.rs:24-95).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <new>
#include <string>
#include <vector>

#include "nsx_device.h"
#include "nsx_internal.h"
#include "nsx_kernels.h"

namespace nsx {

typedef void (*UpdateFn)(StepParams, DevPtrs, int32_t);
hipError_t enqueue_step(nsx_ctx* ctx, int32_t off, hipEvent_t* evs);
hipError_t enqueue_advance(nsx_ctx* ctx, int32_t g);

static thread_local std::string g_thread_err;

nsx_status set_error(nsx_ctx* ctx, nsx_status s, const std::string& msg) {
  if (ctx)
    ctx->err = msg;
  else
    g_thread_err = msg;
  return s;
}

nsx_status hip_error(nsx_ctx* ctx, hipError_t e, const char* what) {
  nsx_status s = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? NSX_E_OOM
                                                                             : NSX_E_HIP;
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) s = NSX_E_NODEVICE;
  return set_error(ctx, s, std::string(what) + ": " + hipGetErrorString(e));
}

nsx_status dev_alloc(nsx_ctx* ctx, void** p, size_t bytes, bool zero) {
  *p = nullptr;
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return set_error(ctx, NSX_E_OOM,
                     "hipMalloc of " + std::to_string(bytes) + " bytes failed: " +
                         hipGetErrorString(e));
  }
  ctx->allocs.push_back(*p);
  if (zero) {
    e = hipMemsetAsync(*p, 0, bytes, ctx->stream);
    if (e != hipSuccess) return hip_error(ctx, e, "hipMemsetAsync");
  }
  return NSX_OK;
}

void dev_free(nsx_ctx* ctx, void* p) {
  if (!p) return;
  auto it = std::find(ctx->allocs.begin(), ctx->allocs.end(), p);
  if (it != ctx->allocs.end()) ctx->allocs.erase(it);
  (void)hipFree(p);
}

// Izhikevich per-neuron c/d with jitter: r ~ U[0,1) from the PARAM stream,
// c = c0 + cj r^2, d = d0 (1 - dj r^2).  Initial v = c (NeuronState::new, mod.rs:45-60), u = 0.
__global__ void k_init_state(StepParams p, DevPtrs s, int model, nsx_config cfg) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.n) return;
  const NeuronId id = neuron_id(p, s, i, i / kBlock);
  s.last[i] = kNeverStep;
  s.pre[i] = TraceRec{0.0f, kNeverStep, 0.0f, kNeverStep};
  s.post[i] = PostRec{0.0f, kNeverStep, kNeverStep, 0u, {}};
  s.count[i] = 0u;
  s.refr[i] = 0;
  s.s0[i] = 0.0f;
  s.s1[i] = 0.0f;
  s.s2[i] = 0.0f;
  if (model == NSX_MODEL_LIF) {
    s.v[i] = p.v_rest;
  } else if (model == NSX_MODEL_IZHIKEVICH) {
    U4 r = draw(id.seed, id.local, 0u, kStreamParam);
    float x = u01(r.x);
    float x2 = x * x;
    const int ty = id.local < id.n_exc ? 0 : 1;
    float cj = cfg.izh_c_jitter * x2;
    float c = cfg.izh_c[ty] + cj;
    float dj = cfg.izh_d_jitter_frac * x2;
    dj = 1.0f - dj;
    float d = cfg.izh_d[ty] * dj;
    s.c[i] = c;
    s.d[i] = d;
    s.v[i] = c;
    s.u[i] = 0.0f;
  } else {
    s.v[i] = -65.0f;
    s.gm[i] = 0.05f;
    s.gh[i] = 0.6f;
    s.gn[i] = 0.32f;
    s.drv[i] = cfg.drive == NSX_DRIVE_OU ? cfg.drive_mean : 0.0f;
  }
}

// The ring slot of step t receives q; the slot counts as live for step t (as if a spike at t-1
// had delivered it).
__global__ void k_stimulate(long long* ring_slot, uint32_t i, long long q, int32_t* ev_last,
                            int32_t t) {
  atomicAdd((unsigned long long*)(ring_slot + i), (unsigned long long)q);
  if (ev_last[0] < t - 1) ev_last[0] = t - 1;
}

static int poisson_table(double lambda, uint32_t* thr, int cap) {
  double pr = exp(-lambda), cdf = pr;
  int k = 0;
  while (k < cap) {
    double t = floor(cdf * 4294967296.0);
    thr[k] = t >= 4294967295.0 ? 0xffffffffu : (uint32_t)t;
    ++k;
    if (thr[k - 1] == 0xffffffffu) break;
    pr = pr * lambda / (double)k;
    cdf += pr;
  }
  return k;
}

static nsx_status validate_config(const nsx_config* c) {
  if (!c) return set_error(nullptr, NSX_E_INVALID, "config is NULL");
  if (c->struct_size != sizeof(nsx_config))
    return set_error(nullptr, NSX_E_INVALID,
                     "nsx_config.struct_size mismatch (header/library version skew)");
  if (c->model > NSX_MODEL_HH_SURROGATE)
    return set_error(nullptr, NSX_E_INVALID, "unknown or unsupported model");
  if (c->model == NSX_MODEL_HH_SURROGATE && c->hh_spike_mode != NSX_HH_SPIKE_CROSSING)
    return set_error(nullptr, NSX_E_INVALID, "the HH surrogate needs the crossing spike mode");
  if (c->n_neurons < 1 || c->n_neurons >= 0xffffffffull)
    return set_error(nullptr, NSX_E_INVALID, "n_neurons must be in [1, 2^32-1)");
  if (c->n_exc > c->n_neurons) return set_error(nullptr, NSX_E_INVALID, "n_exc > n_neurons");
  if (!(c->dt_ms > 0.0)) return set_error(nullptr, NSX_E_INVALID, "dt_ms must be > 0");
  if (c->ring_slots < 2 || c->ring_slots > 256)
    return set_error(nullptr, NSX_E_INVALID, "ring_slots must be in [2, 256]");
  if (c->frac_bits > 40) return set_error(nullptr, NSX_E_INVALID, "frac_bits must be <= 40");
  if (c->v_substeps < 1) return set_error(nullptr, NSX_E_INVALID, "v_substeps must be >= 1");
  if (c->drive > NSX_DRIVE_OU) return set_error(nullptr, NSX_E_INVALID, "unknown drive");
  if (c->plasticity > NSX_PLAST_TRIPLET)
    return set_error(nullptr, NSX_E_INVALID, "unsupported plasticity kind");
  if (c->plasticity != NSX_PLAST_NONE &&
      (!(c->tau_plus_ms > 0) || !(c->tau_minus_ms > 0) ||
       104.0 * std::max(c->tau_plus_ms, c->tau_minus_ms) / c->dt_ms > (double)(1 << 24)))
    return set_error(nullptr, NSX_E_INVALID, "STDP time constants out of range");
  if (c->plasticity == NSX_PLAST_TRIPLET &&
      (!(c->tau_y_ms > 0) || 104.0 * c->tau_y_ms / c->dt_ms > (double)(1 << 24)))
    return set_error(nullptr, NSX_E_INVALID, "triplet tau_y out of range");
  if (c->plasticity == NSX_PLAST_TRIPLET && c->model == NSX_MODEL_HH_SURROGATE)
    return set_error(nullptr, NSX_E_INVALID, "the HH surrogate runs with pair STDP or none");
  if (c->drive == NSX_DRIVE_OU && !(c->drive_tau_ms > 0))
    return set_error(nullptr, NSX_E_INVALID, "drive_tau_ms must be > 0");
  if (c->drive == NSX_DRIVE_POISSON) {
    double lam = (double)c->drive_inputs * c->drive_rate_hz * c->dt_ms / 1000.0;
    uint32_t thr[kMaxPoisson];
    if (!(lam >= 0) || poisson_table(lam, thr, kMaxPoisson) >= kMaxPoisson)
      return set_error(nullptr, NSX_E_INVALID, "Poisson mean per step too large (>~8)");
  }
  if ((c->model == NSX_MODEL_HH || c->model == NSX_MODEL_HH_SURROGATE) &&
      (c->hh_substeps < 1 || c->hh_substeps > 64))
    return set_error(nullptr, NSX_E_INVALID, "hh_substeps must be in [1, 64]");
  if (c->hh_spike_mode == NSX_HH_SPIKE_REFERENCE_RESET && c->hh_substeps > 1)
    return set_error(nullptr, NSX_E_INVALID, "the reference reset rule takes one sub-step per step");
  return NSX_OK;
}

static void fill_params(nsx_ctx* ctx) {
  const nsx_config& c = ctx->cfg;
  StepParams& p = ctx->p;
  memset(&p, 0, sizeof(p));
  const double dt = c.dt_ms;
  p.n = (uint32_t)c.n_neurons;
  p.n_exc = (uint32_t)c.n_exc;
  p.nblk = (uint32_t)((c.n_neurons + kBlock - 1) / kBlock);
  p.prop_split = 4;
  p.D = c.ring_slots;
  p.ch = (uint32_t)ctx->ch;
  p.seed = c.seed;
  p.scale_in = ldexp(1.0, (int)c.frac_bits);
  p.scale_out = ldexp(1.0, -(int)c.frac_bits);
  p.v_rest = c.lif_v_rest;
  p.v_reset = c.lif_v_reset;
  p.v_th = c.lif_v_th;
  p.lif_r = c.lif_r;
  p.lif_k = (float)(dt / c.lif_tau_m);
  p.refr_steps = (int32_t)round((double)c.lif_t_ref_ms / dt);
  p.dec_e = (float)exp(-dt / c.tau_exc_ms);
  p.dec_i = (float)exp(-dt / c.tau_inh_ms);
  p.dec_n = (float)exp(-dt / c.tau_nmda_ms);
  for (int k = 0; k < 2; ++k) {
    p.izh_a[k] = c.izh_a[k];
    p.izh_b[k] = c.izh_b[k];
  }
  p.izh_h = (float)(dt / c.v_substeps);
  p.dtf = (float)dt;
  p.substeps = c.v_substeps;
  const uint32_t hs = c.hh_substeps ? c.hh_substeps : 1u;
  p.hh = HHConst{c.hh_g_na, c.hh_g_k, c.hh_g_l, c.hh_e_na, c.hh_e_k, c.hh_e_l,
                 (float)(dt / hs), (float)(dt / hs / c.hh_c_m),
                 c.hh_spike_mode == NSX_HH_SPIKE_REFERENCE_RESET ? 1u : 0u};
  p.hh_substeps = hs;
  p.hh_spike_v = c.hh_spike_v;
  p.e_exc = c.e_exc;
  p.e_inh = c.e_inh;
  p.gmax_ampa = c.gmax_ampa;
  p.gmax_nmda = c.gmax_nmda;
  p.gmax_gaba = c.gmax_gaba;
  p.mgc = (float)(c.mg_mM / 3.57);
  p.use_nmda = c.mg_mM > 0.0f ? 1u : 0u;
  p.hh_reset_mode = c.hh_spike_mode == NSX_HH_SPIKE_REFERENCE_RESET ? 1u : 0u;
  p.drive_mean = c.drive_mean;
  p.drive_sigma = c.drive_sigma;
  p.drive_weight = c.drive_weight;
  if (c.drive == NSX_DRIVE_OU) {
    double e = exp(-dt / c.drive_tau_ms);
    p.ou_k1 = (float)e;
    p.ou_k2 = (float)(c.drive_mean * (1.0 - e));
    p.ou_k3 = (float)(c.drive_sigma * sqrt(1.0 - e * e));
  }
  if (c.drive == NSX_DRIVE_POISSON)
    p.npthr = poisson_table((double)c.drive_inputs * c.drive_rate_hz * dt / 1000.0, p.pthr,
                            kMaxPoisson);
  p.a_plus = c.a_plus;
  p.a_minus = c.a_minus;
  p.a3 = c.a3_plus;
  p.w_min = c.w_min;
  p.w_max = c.w_max;
  p.step_counts_cap = 1u << 20;
  p.raster_on = c.record_raster ? 1u : 0u;
  p.raster_cap = c.record_raster ? c.raster_capacity : 0;
}

// Lazy trace decay table P[k] = exp(-k dt / tau) rounded to float; entries past the end read 0.
static nsx_status make_trace_table(nsx_ctx* ctx, double tau, float** out, int32_t* len,
                                   std::vector<float>& h) {
  const double dt = ctx->cfg.dt_ms;
  int n = (int)ceil(104.0 * tau / dt) + 2;
  h.assign((size_t)n, 0.0f);
  for (int k = 0; k < n; ++k) h[(size_t)k] = (float)exp(-(double)k * dt / tau);
  nsx_status s = dev_alloc(ctx, (void**)out, (size_t)n * 4, false);
  if (s) return s;
  NSX_HIP(ctx, hipMemcpy(*out, h.data(), (size_t)n * 4, hipMemcpyHostToDevice));
  *len = n;
  return NSX_OK;
}

static nsx_status check_device(nsx_ctx* ctx, int* dev) {
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count < 1) {
    (void)hipGetLastError();
    return set_error(ctx, NSX_E_NODEVICE, "no HIP device visible");
  }
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, 0);
  if (e != hipSuccess) return hip_error(ctx, e, "hipGetDeviceProperties");
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return set_error(ctx, NSX_E_NODEVICE,
                     std::string("device 0 is ") + prop.gcnArchName + ", library built for gfx950");
  *dev = 0;
  return NSX_OK;
}

void destroy_graph(nsx_ctx* ctx) {
  // no replay of the graph may still be in flight when its executable is destroyed
  if (ctx->graph_exec) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
  if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
  ctx->graph_exec = nullptr;
  ctx->graph = nullptr;
}

nsx_status ring_mark(nsx_ctx* ctx, bool never) {
  // live: a spike at t + D - 2 would keep slots t .. t + 2D - 3 live, every slot the host wrote
  const int32_t v = never ? kNeverStep : (int32_t)ctx->t + ctx->D - 2;
  std::vector<int32_t> h(kEvStripes * kEvStride, v);
  NSX_HIP(ctx, hipMemcpyAsync(ctx->d.ev_last, h.data(), h.size() * 4, hipMemcpyHostToDevice,
                              ctx->stream));
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return NSX_OK;
}

}  // namespace nsx

using namespace nsx;

extern "C" {

uint32_t nsx_abi_version(void) { return NSX_ABI_VERSION; }

const char* nsx_status_string(nsx_status s) {
  switch (s) {
    case NSX_OK: return "ok";
    case NSX_E_NODEVICE: return "no device";
    case NSX_E_OOM: return "out of device memory";
    case NSX_E_INVALID: return "invalid argument";
    case NSX_E_BADCSR: return "bad CSR";
    case NSX_E_HIP: return "HIP runtime error";
    case NSX_E_STATE: return "call out of order";
    case NSX_E_OVERFLOW: return "buffer overflow";
  }
  return "unknown status";
}

const char* nsx_last_error(const void* ctx) {
  if (ctx) return static_cast<const nsx_ctx*>(ctx)->err.c_str();
  return g_thread_err.c_str();
}

nsx_status nsx_device_count(int* count) {
  if (!count) return set_error(nullptr, NSX_E_INVALID, "count is NULL");
  *count = 0;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return set_error(nullptr, NSX_E_NODEVICE, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
  }
  int g = 0;
  for (int d = 0; d < n; ++d) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, d) == hipSuccess &&
        strncmp(prop.gcnArchName, "gfx950", 6) == 0)
      ++g;
  }
  *count = g;
  return NSX_OK;
}

void nsx_config_init(nsx_config* c, uint32_t model) {
  if (!c) return;
  memset(c, 0, sizeof(*c));
  c->struct_size = sizeof(nsx_config);
  c->model = model;
  c->n_neurons = 1;
  c->n_exc = 1;
  c->dt_ms = 0.5;  // SimulationEngine::new, engine.rs:96
  c->v_substeps = 1;
  c->ring_slots = 2;
  c->frac_bits = 32;
  c->seed = 42;  // engine.rs:99
  // NeuronModelParams::default (mod.rs:106-117)
  c->lif_v_rest = -60.0f;
  c->lif_v_th = -50.0f;
  c->lif_v_reset = -65.0f;
  c->lif_tau_m = 8.0f;
  c->lif_t_ref_ms = 1.5f;
  c->lif_r = 2.0f;
  // SynapseDynamics (types.rs:15-54)
  c->tau_exc_ms = 2.0f;
  c->tau_inh_ms = 10.0f;
  c->tau_nmda_ms = 100.0f;
  c->e_exc = 0.0f;
  c->e_inh = -70.0f;
  c->gmax_ampa = 1.0f;
  c->gmax_nmda = 0.5f;
  c->gmax_gaba = 1.0f;
  c->mg_mM = 1.0f;
  // IzhikevichNeuron::reset_state (izhikevich.rs:54-70)
  c->izh_a[0] = 0.02f; c->izh_b[0] = 0.2f; c->izh_c[0] = -65.0f; c->izh_d[0] = 8.0f;
  c->izh_a[1] = 0.02f; c->izh_b[1] = 0.25f; c->izh_c[1] = -65.0f; c->izh_d[1] = 2.0f;
  // HodgkinHuxleyNeuron::reset_state (hodgkin_huxley.rs:98-114)
  c->hh_g_na = 120.0f; c->hh_g_k = 36.0f; c->hh_g_l = 0.3f;
  c->hh_e_na = 50.0f; c->hh_e_k = -77.0f; c->hh_e_l = -54.387f; c->hh_c_m = 1.0f;
  c->hh_spike_v = 0.0f;
  c->hh_spike_mode = NSX_HH_SPIKE_REFERENCE_RESET;
  c->hh_substeps = 1;
  // StdpRule::default (plasticity.rs:17-28), TripletStdpRule::default (59-72)
  c->a_plus = 0.01f; c->a_minus = 0.012f;
  c->tau_plus_ms = 20.0f; c->tau_minus_ms = 20.0f;
  c->w_min = 0.0f; c->w_max = 10.0f;
  c->tau_y_ms = 100.0f; c->a3_plus = 0.0003f;
}

nsx_status nsx_create(const nsx_config* cfg, nsx_ctx** out) {
  if (!out) return set_error(nullptr, NSX_E_INVALID, "out is NULL");
  *out = nullptr;
  nsx_status s = validate_config(cfg);
  if (s) return s;
  nsx_ctx* ctx = new (std::nothrow) nsx_ctx();
  if (!ctx) return set_error(nullptr, NSX_E_OOM, "host allocation failed");
  ctx->cfg = *cfg;
  s = check_device(ctx, &ctx->device);
  if (s) {
    set_error(nullptr, s, ctx->err);
    delete ctx;
    return s;
  }
  hipError_t e = hipSetDevice(ctx->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&ctx->ev_call[0]);
  if (e == hipSuccess) e = hipEventCreate(&ctx->ev_call[1]);
  if (e != hipSuccess) {
    s = hip_error(nullptr, e, "stream / event creation");
    nsx_destroy(ctx);
    return s;
  }
  ctx->n = cfg->n_neurons;
  ctx->ch = cfg->model == NSX_MODEL_IZHIKEVICH ? 1 : 2;
  ctx->D = (int)cfg->ring_slots;
  fill_params(ctx);
  const size_t n = ctx->n;
  DevPtrs& d = ctx->d;
  float** f32s[] = {&d.v, &d.u, &d.gm, &d.gh, &d.gn, &d.s0, &d.s1, &d.s2, &d.drv, &d.c, &d.d,
                   };
  for (auto pp : f32s)
    if ((s = dev_alloc(ctx, (void**)pp, n * 4, true))) goto fail;
  if ((s = dev_alloc(ctx, (void**)&d.refr, n * 4, true))) goto fail;
  if ((s = dev_alloc(ctx, (void**)&d.last, n * 4, true))) goto fail;
  if ((s = dev_alloc(ctx, (void**)&d.count, n * 4, true))) goto fail;
  if ((s = dev_alloc(ctx, (void**)&d.pre, n * sizeof(TraceRec), true))) goto fail;
  if ((s = dev_alloc(ctx, (void**)&d.post, n * sizeof(PostRec), true))) goto fail;
  if ((s = dev_alloc(ctx, (void**)&d.ring, (size_t)ctx->D * ctx->ch * n * 8, true))) goto fail;
  if ((s = dev_alloc(ctx, (void**)&d.spikes, (size_t)ctx->p.nblk * kBlock * 4, false))) goto fail;
  if ((s = dev_alloc(ctx, (void**)&d.bcnt, (size_t)ctx->p.nblk * 4, true))) goto fail;
  if ((s = dev_alloc(ctx, (void**)&d.part, (size_t)ctx->p.nblk * kNumPart * 8, true))) goto fail;
  if ((s = dev_alloc(ctx, (void**)&d.nspk, 16, true))) goto fail;
  if ((s = dev_alloc(ctx, (void**)&d.counters, kNumCounters * 8, true))) goto fail;
  if ((s = dev_alloc(ctx, (void**)&d.step_counts, (size_t)ctx->p.step_counts_cap * 4, true)))
    goto fail;
  if ((s = dev_alloc(ctx, (void**)&d.tbase, 64, true))) goto fail;
  if ((s = dev_alloc(ctx, (void**)&d.ev_last, kEvStripes * kEvStride * 4, true))) goto fail;
  if (cfg->record_raster) {
    if ((s = dev_alloc(ctx, (void**)&d.raster_step, cfg->raster_capacity * 4, false))) goto fail;
    if ((s = dev_alloc(ctx, (void**)&d.raster_id, cfg->raster_capacity * 4, false))) goto fail;
  }
  if (cfg->plasticity != NSX_PLAST_NONE) {
    float* pt = nullptr;
    float* mt = nullptr;
    if ((s = make_trace_table(ctx, cfg->tau_plus_ms, &pt, &ctx->p.ptab_n, ctx->ptab_h))) goto fail;
    if ((s = make_trace_table(ctx, cfg->tau_minus_ms, &mt, &ctx->p.mtab_n, ctx->mtab_h))) goto fail;
    d.ptab = pt;
    d.mtab = mt;
  }
  if (cfg->plasticity == NSX_PLAST_TRIPLET) {
    float* yt = nullptr;
    if ((s = make_trace_table(ctx, cfg->tau_y_ms, &yt, &ctx->p.ytab_n, ctx->ytab_h))) goto fail;
    d.ytab = yt;
  }
  if (cfg->model == NSX_MODEL_HH_SURROGATE &&
      (s = dev_alloc(ctx, (void**)&ctx->sur.fb, (size_t)ctx->p.nblk * 8, true)))
    goto fail;
  if ((s = nsx_reset_state(ctx))) goto fail;
  *out = ctx;
  return NSX_OK;
fail:
  set_error(nullptr, s, ctx->err);
  nsx_destroy(ctx);
  return s;
}

void nsx_destroy(nsx_ctx* ctx) {
  if (!ctx) return;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  destroy_graph(ctx);
  for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
  for (hipEvent_t e : ctx->ev_call)
    if (e) (void)hipEventDestroy(e);
  for (void* p : ctx->allocs) (void)hipFree(p);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

nsx_status nsx_reset_state(nsx_ctx* ctx) {
  if (!ctx) return set_error(nullptr, NSX_E_INVALID, "ctx is NULL");
  const uint32_t g = (uint32_t)((ctx->n + 255) / 256);
  hipLaunchKernelGGL(k_init_state, dim3(g), dim3(256), 0, ctx->stream, ctx->p, ctx->d,
                     (int)ctx->cfg.model, ctx->cfg);
  NSX_HIP(ctx, hipGetLastError());
  NSX_HIP(ctx, hipMemsetAsync(ctx->d.ring, 0, (size_t)ctx->D * ctx->ch * ctx->n * 8, ctx->stream));
  NSX_HIP(ctx, hipMemsetAsync(ctx->d.nspk, 0, 16, ctx->stream));
  NSX_HIP(ctx, hipMemsetAsync(ctx->d.bcnt, 0, (size_t)ctx->p.nblk * 4, ctx->stream));
  NSX_HIP(ctx, hipMemsetAsync(ctx->d.part, 0, (size_t)ctx->p.nblk * kNumPart * 8, ctx->stream));
  NSX_HIP(ctx, hipMemsetAsync(ctx->d.step_counts, 0, (size_t)ctx->p.step_counts_cap * 4,
                              ctx->stream));
  NSX_HIP(ctx, hipMemsetAsync(ctx->d.counters, 0, kNumCounters * 8, ctx->stream));
  NSX_HIP(ctx, hipMemsetAsync(ctx->d.tbase, 0, 64, ctx->stream));
  if (ctx->sur.fb)
    NSX_HIP(ctx, hipMemsetAsync(ctx->sur.fb, 0, (size_t)ctx->p.nblk * 8, ctx->stream));
  if (ctx->d.lfp_part)
    NSX_HIP(ctx, hipMemsetAsync(ctx->d.lfp_part, 0, (size_t)ctx->p.nblk * 8, ctx->stream));
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->hs = {};
  ctx->t = 0;
  ctx->tbase_host = 0;
  nsx_status sr = ring_mark(ctx, true);
  if (sr) return sr;
  return bin_clear(ctx);
}


nsx_status nsx_network_size(nsx_ctx* ctx, uint64_t* n, uint64_t* m) {
  if (!ctx) return set_error(nullptr, NSX_E_INVALID, "ctx is NULL");
  if (n) *n = ctx->n;
  if (m) *m = ctx->m;
  return NSX_OK;
}

static nsx_status copy_f(nsx_ctx* ctx, void* dst, const void* src, size_t bytes,
                         hipMemcpyKind kind) {
  if (!dst || !src || bytes == 0) return NSX_OK;
  NSX_HIP(ctx, hipMemcpyAsync(dst, src, bytes, kind, ctx->stream));
  return NSX_OK;
}

// Trace records are packed {value, step} on the device; the state view splits them.
static nsx_status put_traces(nsx_ctx* ctx, TraceRec* dst, const float* v, const int32_t* st) {
  if (!v && !st) return NSX_OK;
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<TraceRec> h(ctx->n);
  NSX_HIP(ctx, hipMemcpy(h.data(), dst, ctx->n * sizeof(TraceRec), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < ctx->n; ++i) {
    if (v) h[i].v = v[i];
    if (st) h[i].s = st[i];
  }
  NSX_HIP(ctx, hipMemcpy(dst, h.data(), ctx->n * sizeof(TraceRec), hipMemcpyHostToDevice));
  return NSX_OK;
}

static nsx_status get_traces(nsx_ctx* ctx, const TraceRec* src, float* v, int32_t* st) {
  if (!v && !st) return NSX_OK;
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<TraceRec> h(ctx->n);
  NSX_HIP(ctx, hipMemcpy(h.data(), src, ctx->n * sizeof(TraceRec), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < ctx->n; ++i) {
    if (v) v[i] = h[i].v;
    if (st) st[i] = h[i].s;
  }
  return NSX_OK;
}

// Post records: uploading reference-view traces leaves no bump pending.
static nsx_status put_post(nsx_ctx* ctx, const float* v, const int32_t* st, const int32_t* last) {
  if (!v && !st) return NSX_OK;
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<PostRec> h(ctx->n);
  NSX_HIP(ctx, hipMemcpy(h.data(), ctx->d.post, ctx->n * sizeof(PostRec), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < ctx->n; ++i) {
    if (v) h[i].v = v[i];
    if (st) h[i].s = st[i];
    h[i].last = last ? last[i] : kNeverStep;
    h[i].npend = 0;  // the uploaded weights carry every depression
  }
  NSX_HIP(ctx, hipMemcpy(ctx->d.post, h.data(), ctx->n * sizeof(PostRec), hipMemcpyHostToDevice));
  NSX_HIP(ctx, hipMemset(ctx->d.bcnt, 0, (size_t)ctx->p.nblk * 4));
  return NSX_OK;
}

// Whether neuron j is excitatory (in a batched population: local index < its network's n_exc).
static bool host_exc(const nsx_ctx* ctx, uint32_t j) {
  if (!ctx->p.batch) return j < ctx->cfg.n_exc;
  for (const BatchNet& b : ctx->nets_h)
    if (j >= b.offset && j - b.offset < b.n) return j - b.offset < b.n_exc;
  return false;
}

// Post records in the reference's view: the bumps of the last step's spikes, which the device
// applies at the start of the next step, are applied here with the same float operations.
static nsx_status get_post(nsx_ctx* ctx, float* v, int32_t* st) {
  if (!v && !st) return NSX_OK;
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const size_t n = ctx->n, nb = ctx->p.nblk;
  std::vector<PostRec> h(n);
  NSX_HIP(ctx, hipMemcpy(h.data(), ctx->d.post, n * sizeof(PostRec), hipMemcpyDeviceToHost));
  if (ctx->cfg.plasticity == NSX_PLAST_PAIR && ctx->t > 0) {
    std::vector<uint32_t> cnt(nb), seg(nb * kBlock);
    NSX_HIP(ctx, hipMemcpy(cnt.data(), ctx->d.bcnt, nb * 4, hipMemcpyDeviceToHost));
    NSX_HIP(ctx, hipMemcpy(seg.data(), ctx->d.spikes, nb * kBlock * 4, hipMemcpyDeviceToHost));
    const int32_t t = (int32_t)ctx->t - 1;
    for (size_t b = 0; b < nb; ++b)
      for (uint32_t k = 0; k < cnt[b]; ++k) {
        const uint32_t j = seg[b * kBlock + k];
        if (!host_exc(ctx, j)) continue;
        const int32_t dk = t - h[j].s;
        const float q = (dk >= 0 && dk < ctx->p.mtab_n) ? ctx->mtab_h[(size_t)dk] : 0.0f;
        float x = h[j].v * q;
        x = x * q;
        h[j].v = x + 1.0f;
        h[j].s = t;
      }
  }
  for (size_t i = 0; i < n; ++i) {
    if (v) v[i] = h[i].v;
    if (st) st[i] = h[i].s;
  }
  return NSX_OK;
}

nsx_status nsx_upload_state(nsx_ctx* ctx, const nsx_state_view* sv) {
  if (!ctx || !sv) return set_error(ctx, NSX_E_INVALID, "NULL argument");
  const size_t b = ctx->n * 4;
  const hipMemcpyKind k = hipMemcpyHostToDevice;
  DevPtrs& d = ctx->d;
  nsx_status s;
  if ((s = copy_f(ctx, d.v, sv->v, b, k)) || (s = copy_f(ctx, d.u, sv->u, b, k)) ||
      (s = copy_f(ctx, d.gm, sv->m, b, k)) || (s = copy_f(ctx, d.gh, sv->h, b, k)) ||
      (s = copy_f(ctx, d.gn, sv->n, b, k)) || (s = copy_f(ctx, d.s0, sv->syn0, b, k)) ||
      (s = copy_f(ctx, d.s1, sv->syn1, b, k)) || (s = copy_f(ctx, d.s2, sv->syn2, b, k)) ||
      (s = copy_f(ctx, d.drv, sv->drive, b, k)) || (s = copy_f(ctx, d.refr, sv->refractory, b, k)) ||
      (s = copy_f(ctx, d.last, sv->last_spike_step, b, k)) ||
      (s = copy_f(ctx, d.count, sv->spike_count, b, k)) ||
      (s = put_traces(ctx, d.pre, sv->pre_trace, sv->pre_step)) ||
      (s = put_post(ctx, sv->post_trace, sv->post_step, sv->last_spike_step)) ||
      (s = copy_f(ctx, d.ring, sv->ring, (size_t)ctx->D * ctx->ch * ctx->n * 8, k)) ||
      (s = copy_f(ctx, d.c, sv->izh_c, b, k)) || (s = copy_f(ctx, d.d, sv->izh_d, b, k)))
    return s;
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if ((s = ring_mark(ctx, false))) return s;
  return bin_clear(ctx);  // the uploaded ring holds every pending input
}

nsx_status nsx_read_state(nsx_ctx* ctx, nsx_state_view* sv) {
  if (!ctx || !sv) return set_error(ctx, NSX_E_INVALID, "NULL argument");
  const size_t b = ctx->n * 4;
  const hipMemcpyKind k = hipMemcpyDeviceToHost;
  DevPtrs& d = ctx->d;
  nsx_status s = bin_materialize(ctx);  // binned events in flight join the ring
  if (s) return s;
  if ((s = copy_f(ctx, sv->v, d.v, b, k)) || (s = copy_f(ctx, sv->u, d.u, b, k)) ||
      (s = copy_f(ctx, sv->m, d.gm, b, k)) || (s = copy_f(ctx, sv->h, d.gh, b, k)) ||
      (s = copy_f(ctx, sv->n, d.gn, b, k)) || (s = copy_f(ctx, sv->syn0, d.s0, b, k)) ||
      (s = copy_f(ctx, sv->syn1, d.s1, b, k)) || (s = copy_f(ctx, sv->syn2, d.s2, b, k)) ||
      (s = copy_f(ctx, sv->drive, d.drv, b, k)) || (s = copy_f(ctx, sv->refractory, d.refr, b, k)) ||
      (s = copy_f(ctx, sv->last_spike_step, d.last, b, k)) ||
      (s = copy_f(ctx, sv->spike_count, d.count, b, k)) ||
      (s = get_traces(ctx, d.pre, sv->pre_trace, sv->pre_step)) ||
      (s = get_post(ctx, sv->post_trace, sv->post_step)) ||
      (s = copy_f(ctx, sv->ring, d.ring, (size_t)ctx->D * ctx->ch * ctx->n * 8, k)) ||
      (s = copy_f(ctx, sv->izh_c, d.c, b, k)) || (s = copy_f(ctx, sv->izh_d, d.d, b, k)))
    return s;
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return NSX_OK;
}

nsx_status nsx_read_weights(nsx_ctx* ctx, float* weight, uint64_t m) {
  if (!ctx || !weight) return set_error(ctx, NSX_E_INVALID, "NULL argument");
  if (!ctx->has_network) return set_error(ctx, NSX_E_STATE, "no network loaded");
  if (m < ctx->m) return set_error(ctx, NSX_E_INVALID, "weight buffer too small");
  nsx_status s = flush_pending(ctx, true);
  if (s) return s;
  NSX_HIP(ctx, hipMemcpyAsync(weight, ctx->w, ctx->m * 4, hipMemcpyDeviceToHost, ctx->stream));
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return NSX_OK;
}

nsx_status nsx_upload_weights(nsx_ctx* ctx, const float* weight, uint64_t m) {
  if (!ctx || (!weight && m)) return set_error(ctx, NSX_E_INVALID, "NULL argument");
  if (!ctx->has_network) return set_error(ctx, NSX_E_STATE, "no network loaded");
  if (m != ctx->m) return set_error(ctx, NSX_E_INVALID, "weight count differs from the network's");
  nsx_status s = flush_pending(ctx, false);  // depressions pending on the old weights
  if (s) return s;
  if (ctx->pull_built) {
    // pull delivery: input owed by past spikes joins the ring with the weights it left with;
    // the re-sorted copy is rebuilt with the new weights before the next step
    if ((s = bin_materialize(ctx))) return s;
    pull_release(ctx);
    if (ctx->delivery == NSX_DELIVERY_PULL) ctx->bin_dirty = true;
    destroy_graph(ctx);
  }
  if (m)
    NSX_HIP(ctx, hipMemcpyAsync(ctx->w, weight, m * 4, hipMemcpyHostToDevice, ctx->stream));
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return NSX_OK;
}

nsx_status nsx_stimulate(nsx_ctx* ctx, uint64_t neuron, float current) {
  if (!ctx) return set_error(nullptr, NSX_E_INVALID, "ctx is NULL");
  if (neuron >= ctx->n) return NSX_OK;  // SimulationEngine::stimulate ignores out-of-range ids
  const uint32_t slot = (uint32_t)(ctx->t % (uint64_t)ctx->D);
  long long q = (long long)llrint((double)current * ctx->p.scale_in);
  hipLaunchKernelGGL(k_stimulate, dim3(1), dim3(1), 0, ctx->stream,
                     ctx->d.ring + (size_t)slot * ctx->ch * ctx->n, (uint32_t)neuron, q,
                     ctx->d.ev_last, (int32_t)ctx->t);
  NSX_HIP(ctx, hipGetLastError());
  return NSX_OK;
}

nsx_status nsx_get_step(nsx_ctx* ctx, uint64_t* step) {
  if (!ctx || !step) return set_error(ctx, NSX_E_INVALID, "NULL argument");
  *step = ctx->t;
  return NSX_OK;
}

nsx_status nsx_set_step(nsx_ctx* ctx, uint64_t step) {
  if (!ctx) return set_error(nullptr, NSX_E_INVALID, "ctx is NULL");
  if (step >= (1ull << 30)) return set_error(ctx, NSX_E_INVALID, "step index limit");
  // deferred depressions are applied, then the post trace bumps pending from the last step
  // (keyed by that step)
  if (ctx->has_network) {
    nsx_status sf = flush_pending(ctx, true);
    if (!sf) sf = bin_materialize(ctx);  // the ring then holds every pending input
    if (sf) return sf;
  }
  if (ctx->cfg.plasticity == NSX_PLAST_PAIR && ctx->t > 0) {
    std::vector<float> v(ctx->n);
    std::vector<int32_t> st(ctx->n), last(ctx->n);
    nsx_status s = get_post(ctx, v.data(), st.data());
    if (s) return s;
    NSX_HIP(ctx, hipMemcpy(last.data(), ctx->d.last, ctx->n * 4, hipMemcpyDeviceToHost));
    if ((s = put_post(ctx, v.data(), st.data(), last.data()))) return s;
  }
  NSX_HIP(ctx, hipMemsetAsync(ctx->d.nspk, 0, 16, ctx->stream));
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->t = step;
  return ring_mark(ctx, false);  // the ring's slots are read at the new step indices
}

nsx_status nsx_set_prop_split(nsx_ctx* ctx, uint32_t waves_per_block) {
  if (!ctx) return set_error(nullptr, NSX_E_INVALID, "ctx is NULL");
  if (waves_per_block < 1 || waves_per_block > 64)
    return set_error(ctx, NSX_E_INVALID, "waves_per_block must be in [1, 64]");
  ctx->p.prop_split = waves_per_block;
  destroy_graph(ctx);  // captured launches hold the old shape
  return NSX_OK;
}

nsx_status nsx_set_update_width(nsx_ctx* ctx, uint32_t neurons_per_thread) {
  if (!ctx) return set_error(nullptr, NSX_E_INVALID, "ctx is NULL");
  if (neurons_per_thread != 0 && neurons_per_thread != 1 && neurons_per_thread != 4)
    return set_error(ctx, NSX_E_INVALID, "neurons_per_thread must be 0, 1 or 4");
  ctx->update_width = neurons_per_thread;
  destroy_graph(ctx);  // captured launches hold the old kernel
  return NSX_OK;
}

nsx_status nsx_set_row_split(nsx_ctx* ctx, uint32_t slots) {
  if (!ctx) return set_error(nullptr, NSX_E_INVALID, "ctx is NULL");
  if (slots != 0 && (slots < 16 || slots > (1u << 20)))
    return set_error(ctx, NSX_E_INVALID, "row split threshold must be 0 or 16 .. 2^20 slots");
  ctx->p.row_split = slots;
  destroy_graph(ctx);  // captured launches hold the old parameters
  return NSX_OK;
}

nsx_status nsx_set_syn_block(nsx_ctx* ctx, uint32_t threads) {
  if (!ctx) return set_error(nullptr, NSX_E_INVALID, "ctx is NULL");
  if (threads != 64 && threads != 128 && threads != 256 && threads != 512)
    return set_error(ctx, NSX_E_INVALID, "synapse workgroup size must be 64, 128, 256 or 512");
  ctx->syn_block = threads;
  destroy_graph(ctx);
  return NSX_OK;
}

nsx_status nsx_set_persistent(nsx_ctx* ctx, int enable) {
  if (!ctx) return set_error(nullptr, NSX_E_INVALID, "ctx is NULL");
  ctx->persistent = enable != 0;
  destroy_graph(ctx);
  return NSX_OK;
}

nsx_status nsx_get_persistent(nsx_ctx* ctx, uint32_t* requested, uint32_t* applies) {
  if (!ctx) return set_error(nullptr, NSX_E_INVALID, "ctx is NULL");
  if (applies && ctx->has_network) {
    nsx_status s = epoch_prepare(ctx);  // whether epochs take the network over
    if (s) return s;
  }
  if (requested) *requested = ctx->persistent ? 1u : 0u;
  if (applies) *applies = ctx->has_network && persist_applies(ctx) ? 1u : 0u;
  return NSX_OK;
}

nsx_status nsx_set_epoch(nsx_ctx* ctx, uint32_t max_steps) {
  if (!ctx) return set_error(nullptr, NSX_E_INVALID, "ctx is NULL");
  if (max_steps > 1024) return set_error(ctx, NSX_E_INVALID, "epoch longer than 1024 steps");
  ctx->epoch_cap = max_steps;
  destroy_graph(ctx);  // captured launches hold the old grouping
  return NSX_OK;
}

nsx_status nsx_get_epoch(nsx_ctx* ctx, uint32_t* steps) {
  if (!ctx || !steps) return set_error(ctx, NSX_E_INVALID, "NULL argument");
  nsx_status s = epoch_prepare(ctx);
  if (s) return s;
  *steps = epoch_len(ctx);
  return NSX_OK;
}

nsx_status nsx_set_graph_steps(nsx_ctx* ctx, uint32_t steps_per_graph) {
  if (!ctx) return set_error(nullptr, NSX_E_INVALID, "ctx is NULL");
  if (steps_per_graph > 4096) return set_error(ctx, NSX_E_INVALID, "steps_per_graph > 4096");
  destroy_graph(ctx);
  ctx->graph_steps = steps_per_graph;
  return NSX_OK;
}

nsx_status nsx_set_profiling(nsx_ctx* ctx, int enable) {
  if (!ctx) return set_error(nullptr, NSX_E_INVALID, "ctx is NULL");
  // switching on starts a fresh measurement; switching off keeps it for nsx_kernel_time
  if (enable && !ctx->profiling)
    for (int k = 0; k < 4; ++k) {
      ctx->kt_ms[k] = 0;
      ctx->kt_launches[k] = 0;
    }
  ctx->profiling = enable != 0;
  return NSX_OK;
}

nsx_status nsx_kernel_time(nsx_ctx* ctx, int kind, double* total_ms, uint64_t* launches,
                           uint64_t* bytes) {
  if (!ctx || kind < 0 || kind > 3) return set_error(ctx, NSX_E_INVALID, "bad argument");
  if (total_ms) *total_ms = ctx->kt_ms[kind];
  if (launches) *launches = ctx->kt_launches[kind];
  // slot counts of the last nsx_step call: kind 1 -> potentiated slots, kind 2 -> depressed
  if (bytes) *bytes = kind == 1 ? ctx->part_delta[kPartLtp] : (kind == 2 ? ctx->part_delta[kPartLtd] : 0);
  return NSX_OK;
}

// Totals of the per-block counters (spikes, events, potentiated and depressed slots).
static nsx_status read_totals(nsx_ctx* ctx, uint64_t tot[kNumPart]) {
  const size_t nb = ctx->p.nblk;
  ctx->part_host.resize(nb * kNumPart);
  NSX_HIP(ctx, hipMemcpyAsync(ctx->part_host.data(), ctx->d.part, nb * kNumPart * 8,
                              hipMemcpyDeviceToHost, ctx->stream));
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int q = 0; q < kNumPart; ++q) {
    uint64_t x = 0;
    for (size_t b = 0; b < nb; ++b) x += ctx->part_host[q * nb + b];
    tot[q] = x;
  }
  return NSX_OK;
}

// Capture G steps (+ the tbase advance) into one graph.  Per-kernel timing needs events on the
// stream (HIP does not time events recorded by graph nodes), so profiled steps launch directly.
static nsx_status capture_graph(nsx_ctx* ctx) {
  destroy_graph(ctx);
  const int32_t G = (int32_t)ctx->graph_steps;
  const int32_t E = (int32_t)epoch_len(ctx);
  NSX_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = hipSuccess;
  for (int32_t k = 0; k < G && e == hipSuccess;) {
    if (E > 1) {
      const int32_t L = std::min(E, G - k);
      e = enqueue_epoch(ctx, k, (uint32_t)L, nullptr);
      k += L;
    } else {
      e = enqueue_step(ctx, k, nullptr);
      k += 1;
    }
  }
  if (e == hipSuccess) e = enqueue_advance(ctx, G);
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
  if (e != hipSuccess) return hip_error(ctx, e, "graph capture");
  if (e2 != hipSuccess) return hip_error(ctx, e2, "hipStreamEndCapture");
  ctx->graph = g;
  NSX_HIP(ctx, hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0));
  return NSX_OK;
}

nsx_status nsx_step(nsx_ctx* ctx, uint64_t n_steps, nsx_step_stats* st) {
  if (!ctx) return set_error(nullptr, NSX_E_INVALID, "ctx is NULL");
  if (!ctx->has_network) return set_error(ctx, NSX_E_STATE, "no network loaded");
  if (ctx->t + n_steps >= (1ull << 30)) return set_error(ctx, NSX_E_INVALID, "step index limit");
  if (ctx->cfg.model == NSX_MODEL_HH_SURROGATE && !(ctx->sur.has_params && ctx->sur.has_norm))
    return set_error(ctx, NSX_E_STATE, "surrogate parameters not set (train or load them)");
  if (ctx->bin_dirty) {  // binned delivery sized for the network now loaded
    nsx_status sb = bin_configure(ctx, ctx->delivery, ctx->bin_cap_req, ctx->bin_te_req);
    if (sb) {  // a network binned delivery cannot take: the engine stays atomic
      ctx->delivery = NSX_DELIVERY_ATOMIC;
      ctx->bin_dirty = false;
      return sb;
    }
  }
  {
    nsx_status se = epoch_prepare(ctx);  // the network's shortest delay, epoch segments
    if (se) return se;
  }
  if (ctx->p.stats_on && n_steps > ctx->p.stat_cap) {
    // the per-step count ring holds stat_cap steps: run in chunks, folding after each
    nsx_step_stats one{};
    uint64_t left = n_steps, spikes = 0, events = 0;
    double wall = 0.0;
    while (left) {
      const uint64_t k = std::min<uint64_t>(left, ctx->p.stat_cap);
      nsx_status s = nsx_step(ctx, k, &one);
      if (s) return s;
      spikes += one.spikes;
      events += one.events;
      wall += one.wall_ms;
      left -= k;
    }
    if (st) {
      *st = one;
      st->steps_done = n_steps;
      st->spikes = spikes;
      st->events = events;
      st->wall_ms = wall;
    }
    return NSX_OK;
  }
  if (ctx->p.stats_on && n_steps)
    NSX_HIP(ctx, hipMemsetAsync(ctx->d.stat_counts, 0,
                                (size_t)ctx->p.stat_cap * kStatStripes * 4, ctx->stream));
  uint64_t c0[kNumPart];
  nsx_status s0 = read_totals(ctx, c0);
  if (s0) return s0;
  // the context's two call events: created once, so no return path below leaks them
  hipEvent_t e_beg = ctx->ev_call[0], e_end = ctx->ev_call[1];
  NSX_HIP(ctx, hipEventRecord(e_beg, ctx->stream));
  const bool prof = ctx->profiling;
  const size_t nev = 5;  // events around the update, counting, synapse and arrival launches
  if (prof && ctx->ev_pool.size() < nev * n_steps) {
    size_t have = ctx->ev_pool.size();
    ctx->ev_pool.resize(nev * n_steps);
    for (size_t k = have; k < ctx->ev_pool.size(); ++k)
      NSX_HIP(ctx, hipEventCreate(&ctx->ev_pool[k]));
  }
  uint64_t done = 0;
  const bool persist = persist_applies(ctx) && !prof;
  if (persist) {
    // windows of graph_steps steps (1024 when graphs are off), one cooperative launch each
    if (!ctx->d_bar) {
      nsx_status sa = dev_alloc(ctx, (void**)&ctx->d_bar, 64, true);
      if (sa) return sa;
    }
    const uint64_t W = ctx->graph_steps > 0 ? ctx->graph_steps : 1024;
    while (done < n_steps) {
      const uint32_t L = (uint32_t)std::min<uint64_t>(W, n_steps - done);
      const int32_t off = (int32_t)(ctx->t - (uint64_t)ctx->tbase_host);
      const hipError_t e = launch_persist(ctx, off, L);
      if (e != hipSuccess) return hip_error(ctx, e, "persistent launch");
      ctx->t += L;
      done += L;
    }
  } else if (ctx->graph_steps > 0 && !prof) {
    const uint64_t G = ctx->graph_steps;
    if (n_steps >= G) {
      // graph offsets are relative to tbase: move tbase to the current step first
      if ((int64_t)ctx->tbase_host != (int64_t)ctx->t) {
        int32_t tb = (int32_t)ctx->t;
        NSX_HIP(ctx, hipMemcpyAsync(ctx->d.tbase, &tb, 4, hipMemcpyHostToDevice, ctx->stream));
        NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        ctx->tbase_host = tb;
      }
      if (!ctx->graph_exec) {
        nsx_status s = capture_graph(ctx);
        if (s) return s;
      }
      while (n_steps - done >= G) {
        NSX_HIP(ctx, hipGraphLaunch(ctx->graph_exec, ctx->stream));
        done += G;
        ctx->t += G;
        ctx->tbase_host += (int32_t)G;
      }
    }
  }
  // direct launches: one step, or one epoch of up to E steps, per launch group
  const uint64_t E = epoch_len(ctx);
  uint64_t groups = 0;
  for (; done < n_steps; ++groups) {
    hipEvent_t* evs = prof ? &ctx->ev_pool[nev * groups] : nullptr;
    const int32_t off = (int32_t)(ctx->t - (uint64_t)ctx->tbase_host);
    const uint64_t L = E > 1 ? std::min<uint64_t>(E, n_steps - done) : 1;
    hipError_t e = E > 1 ? enqueue_epoch(ctx, off, (uint32_t)L, evs) : enqueue_step(ctx, off, evs);
    if (e != hipSuccess) return hip_error(ctx, e, "step launch");
    ctx->t += L;
    done += L;
  }
  NSX_HIP(ctx, hipEventRecord(e_end, ctx->stream));
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (persist) {
    uint32_t gave_up = 0;
    NSX_HIP(ctx, hipMemcpy(&gave_up, ctx->d_bar + 1, 4, hipMemcpyDeviceToHost));
    if (gave_up)
      return set_error(ctx, NSX_E_HIP,
                       "persistent launch: a grid barrier gave up waiting; the state after this "
                       "call is not valid");
  }
  uint64_t c1[kNumPart];
  nsx_status s1 = read_totals(ctx, c1);
  if (s1) return s1;
  unsigned long long rast = 0;
  if (ctx->cfg.record_raster)
    NSX_HIP(ctx, hipMemcpy(&rast, ctx->d.counters + kCntRaster, 8, hipMemcpyDeviceToHost));
  float wall = 0.0f;
  NSX_HIP(ctx, hipEventElapsedTime(&wall, e_beg, e_end));
  if (prof) {
    for (uint64_t k = 0; k < groups; ++k) {
      hipEvent_t* evs = &ctx->ev_pool[nev * k];
      // event interval -> nsx_kernel_time kind: binned delivery runs update, synapse pass,
      // bucket passes, arrival; the others update, (none), synapse pass (or pull), (none)
      static const int kind_bin[4] = {0, 1, 2, 3}, kind_other[4] = {0, 2, 1, 3};
      const int* kind_of = ctx->p.bin_on ? kind_bin : kind_other;
      for (size_t q = 0; q + 1 < nev; ++q) {
        const int kd = kind_of[q];
        if ((kd == 2 || kd == 3) && !ctx->p.bin_on) continue;
        float ms = 0.0f;
        NSX_HIP(ctx, hipEventElapsedTime(&ms, evs[q], evs[q + 1]));
        ctx->kt_ms[kd] += ms;
        ctx->kt_launches[kd] += 1;
      }
    }
  }
  ctx->last_call_steps = (uint32_t)std::min<uint64_t>(n_steps, ctx->p.step_counts_cap);
  ctx->last_call_first_step = ctx->t - n_steps;
  if (rast > ctx->p.raster_cap && ctx->cfg.record_raster)
    set_error(ctx, NSX_E_OVERFLOW, "raster capacity exceeded; oldest spikes kept");
  for (int q = 0; q < kNumPart; ++q) ctx->part_delta[q] = c1[q] - c0[q];
  if (ctx->p.stats_on) {
    nsx_status sf = stats_fold(ctx, ctx->t - n_steps, n_steps);
    if (sf) return sf;
  }
  if (st) {
    st->steps_done = n_steps;
    st->step_index = ctx->t;
    st->spikes = c1[kPartSpikes] - c0[kPartSpikes];
    st->events = c1[kPartEvents] - c0[kPartEvents];
    st->total_spikes = c1[kPartSpikes];
    st->total_events = c1[kPartEvents];
    st->sim_time_ms = (double)ctx->t * ctx->cfg.dt_ms;
    st->wall_ms = wall;
  }
  return NSX_OK;
}

nsx_status nsx_simulate_ms(nsx_ctx* ctx, double duration_ms, nsx_step_stats* st) {
  if (!ctx) return set_error(nullptr, NSX_E_INVALID, "ctx is NULL");
  // SimulationEngine::simulate_ms: steps = (duration_ms / dt) as usize (engine.rs:792-797)
  double q = duration_ms / ctx->cfg.dt_ms;
  uint64_t steps = q > 0 ? (uint64_t)q : 0;
  return nsx_step(ctx, steps, st);
}

nsx_status nsx_read_raster(nsx_ctx* ctx, uint32_t* steps, uint32_t* neurons, uint64_t cap,
                           uint64_t* n_out, int clear) {
  if (!ctx) return set_error(nullptr, NSX_E_INVALID, "ctx is NULL");
  if (!ctx->cfg.record_raster) return set_error(ctx, NSX_E_STATE, "raster recording is off");
  unsigned long long cnt = 0;
  NSX_HIP(ctx, hipMemcpy(&cnt, ctx->d.counters + kCntRaster, 8, hipMemcpyDeviceToHost));
  uint64_t avail = std::min<uint64_t>(cnt, ctx->p.raster_cap);
  if (n_out) *n_out = avail;
  if (steps && neurons && cap) {
    std::vector<uint32_t> hs(avail), hi(avail);
    if (avail) {
      NSX_HIP(ctx, hipMemcpy(hs.data(), ctx->d.raster_step, avail * 4, hipMemcpyDeviceToHost));
      NSX_HIP(ctx, hipMemcpy(hi.data(), ctx->d.raster_id, avail * 4, hipMemcpyDeviceToHost));
    }
    std::vector<uint64_t> key(avail);
    for (uint64_t k = 0; k < avail; ++k) key[k] = ((uint64_t)hs[k] << 32) | hi[k];
    std::sort(key.begin(), key.end());
    uint64_t nc = std::min<uint64_t>(cap, avail);
    for (uint64_t k = 0; k < nc; ++k) {
      steps[k] = (uint32_t)(key[k] >> 32);
      neurons[k] = (uint32_t)key[k];
    }
  }
  if (clear) NSX_HIP(ctx, hipMemset(ctx->d.counters + kCntRaster, 0, 8));
  if (cnt > ctx->p.raster_cap)
    return set_error(ctx, NSX_E_OVERFLOW, "raster capacity exceeded");
  return NSX_OK;
}

nsx_status nsx_read_step_counts(nsx_ctx* ctx, uint32_t* counts, uint64_t cap, uint64_t* n_out) {
  if (!ctx) return set_error(nullptr, NSX_E_INVALID, "ctx is NULL");
  if (!ctx->cfg.record_raster)
    return set_error(ctx, NSX_E_STATE, "per-step counts are recorded with the raster only");
  const uint64_t n = std::min<uint64_t>(ctx->last_call_steps, cap);
  if (n_out) *n_out = ctx->last_call_steps;
  const uint32_t C = ctx->p.step_counts_cap;
  uint64_t k = 0;
  while (k < n) {  // at most two contiguous pieces of the circular buffer
    const uint64_t t = ctx->last_call_first_step + k;
    const uint64_t piece = std::min<uint64_t>(n - k, C - (t % C));
    NSX_HIP(ctx, hipMemcpy(counts + k, ctx->d.step_counts + (t % C), piece * 4,
                           hipMemcpyDeviceToHost));
    k += piece;
  }
  return NSX_OK;
}

}  // extern "C"
