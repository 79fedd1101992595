// nsx_backend.hip -- drop-in for GpuBackend (src/simulation/gpu_backend.rs:47-1037).
//
// Same methods, argument meaning and per-call host round trip as the wgpu backend, so
// SimulationEngine::init_gpu / step_gpu (engine.rs:182-252, 700-790) can bind them unchanged.
// Differences, each deliberate:
//   * LIF refractory semantics follow LifNeuron::step (lif.rs:42-45: V frozen), not
//     lif_update.wgsl (V := resting).
//   * HH runs (the wgpu path panics, gpu_backend.rs:838-849) with HodgkinHuxleyNeuron::step
//     semantics; params6 = [g_na, g_k, g_l, e_na, e_k, e_l], C_m = 1.
//   * step_spmv sums each target's input as int64 fixed point (2^-32 mV), so the result does
//     not depend on the order spikes are processed (spmv.wgsl's float CAS loop does).
#include <math.h>
#include <string.h>

#include <new>
#include <string>

#include "nsx_device.h"
#include "nsx_internal.h"

namespace nsx {

__device__ inline float hh_ratio_b(float x, float scale, float k) {
  if (fabsf(x) < 1e-3f) {
    float y = x * 0.05f;
    y = 1.0f + y;
    return k * y;
  }
  float e = expf_(x * -0.1f);
  float den = 1.0f - e;
  float num = scale * x;
  return num / den;
}

struct BackendPtrs {
  float *v, *u, *t_last, *I, *params, *gm, *gh, *gn;
  int32_t* refr;
  uint32_t *count, *spiked, *list, *nlist;
  const uint32_t *row_ptr, *col;
  const float* w;
  long long* acc;
};

__global__ void k_backend_neurons(BackendPtrs b, uint32_t n, uint32_t model, float dt,
                                  float sim_time) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  bool spiked = false;
  if (i < n) {
    const float* pr = b.params + (size_t)i * 6;
    const float I = b.I[i];
    float v = b.v[i];
    if (model == 0) {  // LifNeuron::step, lif.rs:15-61
      int32_t refr = b.refr[i];
      if (refr > 0) {
        b.refr[i] = refr - 1;
      } else {
        float x = v - pr[0];
        x = -x;
        float y = pr[5] * I;
        x = x + y;
        x = x * dt;
        x = x / pr[3];
        v = v + x;
        if (v >= pr[1]) {
          v = pr[2];
          float q = pr[4] / dt;
          b.refr[i] = (int32_t)roundf(q);
          spiked = true;
        }
        b.v[i] = v;
      }
    } else if (model == 1) {  // IzhikevichNeuron::step, izhikevich.rs:20-52
      float u = b.u[i];
      if (v >= 30.0f) {
        v = pr[2];
        u = u + pr[3];
        spiked = true;
      } else {
        float t1 = 0.04f * v;
        t1 = t1 * v;
        float t2 = 5.0f * v;
        float acc = t1 + t2;
        acc = acc + 140.0f;
        acc = acc - u;
        acc = acc + I;
        acc = acc * dt;
        float t3 = pr[1] * v;
        t3 = t3 - u;
        t3 = pr[0] * t3;
        t3 = t3 * dt;
        v = v + acc;
        u = u + t3;
      }
      b.v[i] = v;
      b.u[i] = u;
    } else {  // HodgkinHuxleyNeuron::step, hodgkin_huxley.rs:46-96
      float am = hh_ratio_b(v + 40.0f, 0.1f, 1.0f);
      float bm = 4.0f * expf_((v + 65.0f) * (-1.0f / 18.0f));
      float ah = 0.07f * expf_((v + 65.0f) * -0.05f);
      float eh = expf_((v + 35.0f) * -0.1f);
      eh = 1.0f + eh;
      float bh = 1.0f / eh;
      float an = hh_ratio_b(v + 55.0f, 0.01f, 0.1f);
      float bn = 0.125f * expf_((v + 65.0f) * -0.0125f);
      float m = b.gm[i], h = b.gh[i], nn = b.gn[i];
      float t1, t2;
      t1 = am * (1.0f - m); t2 = bm * m; t1 = t1 - t2; t1 = t1 * dt; m = m + t1;
      t1 = ah * (1.0f - h); t2 = bh * h; t1 = t1 - t2; t1 = t1 * dt; h = h + t1;
      t1 = an * (1.0f - nn); t2 = bn * nn; t1 = t1 - t2; t1 = t1 * dt; nn = nn + t1;
      float m3 = m * m;
      m3 = m3 * m;
      float ina = pr[0] * m3;
      ina = ina * h;
      ina = ina * (pr[3] - v);
      float n2 = nn * nn;
      float n4 = n2 * n2;
      float ik = pr[1] * n4;
      ik = ik * (pr[4] - v);
      float il = pr[2] * (pr[5] - v);
      float acc = I - ina;
      acc = acc - ik;
      acc = acc - il;
      acc = acc * dt;
      v = v + acc;
      if (v >= -55.0f) {
        v = -70.0f;
        spiked = true;
      }
      b.v[i] = v;
      b.gm[i] = m;
      b.gh[i] = h;
      b.gn[i] = nn;
    }
    b.spiked[i] = spiked ? 1u : 0u;
    if (spiked) {
      b.count[i] += 1u;
      b.t_last[i] = sim_time;
    }
  }
  const unsigned long long mask = __ballot(spiked);
  if (mask) {
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(b.nlist, (unsigned)__popcll(mask));
    base = __shfl(base, leader);
    if (spiked) b.list[base + (unsigned)__popcll(mask & ((1ull << lane) - 1ull))] = i;
  }
}

__global__ void k_backend_spmv(BackendPtrs b, double scale) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
  const uint32_t ns = *b.nlist;
  for (uint32_t r = wave; r < ns; r += nwaves) {
    const uint32_t i = b.list[r];
    for (uint32_t k = b.row_ptr[i] + lane; k < b.row_ptr[i + 1]; k += 64)
      atomicAdd((unsigned long long*)&b.acc[b.col[k]], (unsigned long long)quantize(b.w[k], scale));
  }
}

__global__ void k_backend_convert(BackendPtrs b, uint32_t n, double inv_scale) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *b.nlist = 0u;
  if (i >= n) return;
  float x = dequantize(b.acc[i], inv_scale);
  b.acc[i] = 0;
  b.I[i] = b.I[i] + x;
}

static BackendPtrs backend_ptrs(nsx_ctx* ctx) {
  BackendPtrs b;
  b.v = ctx->d.v;
  b.u = ctx->d.u;
  b.t_last = ctx->d.s2;  // last_spike_time (f32) in the backend layout
  b.I = ctx->b_input;
  b.params = ctx->b_params;
  b.gm = ctx->d.gm;
  b.gh = ctx->d.gh;
  b.gn = ctx->d.gn;
  b.refr = ctx->d.refr;
  b.count = ctx->d.count;
  b.spiked = ctx->b_spiked;
  b.list = ctx->d.spikes;
  b.nlist = ctx->d.nspk;
  b.row_ptr = ctx->b_row_ptr32;
  b.col = ctx->col;
  b.w = ctx->w;
  b.acc = ctx->d.ring;
  return b;
}

}  // namespace nsx

using namespace nsx;

static nsx_status need_backend(nsx_ctx* ctx) {
  if (!ctx) return set_error(nullptr, NSX_E_INVALID, "ctx is NULL");
  if (!ctx->backend_mode) return set_error(ctx, NSX_E_STATE, "not a GpuBackend context");
  return NSX_OK;
}

extern "C" {

nsx_status nsx_backend_new(uint64_t num_neurons, uint64_t num_edges, nsx_ctx** out) {
  if (!out) return set_error(nullptr, NSX_E_INVALID, "out is NULL");
  *out = nullptr;
  if (num_neurons < 1 || num_neurons >= 0xffffffffull || num_edges >= 0xffffffffull)
    return set_error(nullptr, NSX_E_INVALID, "sizes out of the u32 range GpuBackend uses");
  nsx_config cfg;
  nsx_config_init(&cfg, NSX_MODEL_LIF);
  cfg.n_neurons = num_neurons;
  cfg.n_exc = num_neurons;
  cfg.ring_slots = 2;
  nsx_ctx* ctx = nullptr;
  nsx_status s = nsx_create(&cfg, &ctx);
  if (s) return s;
  ctx->backend_mode = true;
  const size_t n = num_neurons;
  if ((s = dev_alloc(ctx, (void**)&ctx->b_input, n * 4, true)) ||
      (s = dev_alloc(ctx, (void**)&ctx->b_params, n * 24, true)) ||
      (s = dev_alloc(ctx, (void**)&ctx->b_spiked, n * 4, true)) ||
      (s = dev_alloc(ctx, (void**)&ctx->b_row_ptr32, (n + 1) * 4, true)) ||
      (s = dev_alloc(ctx, (void**)&ctx->col, num_edges * 4, true)) ||
      (s = dev_alloc(ctx, (void**)&ctx->w, num_edges * 4, true))) {
    set_error(nullptr, s, ctx->err);
    nsx_destroy(ctx);
    return s;
  }
  ctx->m = num_edges;
  // NeuronArray::new defaults for HH gates (mod.rs:158-160)
  std::vector<float> g(n);
  std::fill(g.begin(), g.end(), 0.05f);
  (void)hipMemcpy(ctx->d.gm, g.data(), n * 4, hipMemcpyHostToDevice);
  std::fill(g.begin(), g.end(), 0.6f);
  (void)hipMemcpy(ctx->d.gh, g.data(), n * 4, hipMemcpyHostToDevice);
  std::fill(g.begin(), g.end(), 0.32f);
  (void)hipMemcpy(ctx->d.gn, g.data(), n * 4, hipMemcpyHostToDevice);
  (void)hipMemset(ctx->d.ring, 0, n * 8);
  if (hipDeviceSynchronize() != hipSuccess) {
    s = set_error(nullptr, NSX_E_HIP, "backend initialisation failed");
    nsx_destroy(ctx);
    return s;
  }
  *out = ctx;
  return NSX_OK;
}

nsx_status nsx_backend_upload_initial_state(nsx_ctx* ctx, const float* v, const float* u,
                                            const int32_t* refr, const float* t_last,
                                            const uint32_t* count, const float* params6,
                                            const uint32_t* is_output) {
  nsx_status s = need_backend(ctx);
  if (s) return s;
  (void)is_output;  // read back by the host engine only (engine.rs:758)
  const size_t b = ctx->n * 4;
  if (v) NSX_HIP(ctx, hipMemcpy(ctx->d.v, v, b, hipMemcpyHostToDevice));
  if (u) NSX_HIP(ctx, hipMemcpy(ctx->d.u, u, b, hipMemcpyHostToDevice));
  if (refr) NSX_HIP(ctx, hipMemcpy(ctx->d.refr, refr, b, hipMemcpyHostToDevice));
  if (t_last) NSX_HIP(ctx, hipMemcpy(ctx->d.s2, t_last, b, hipMemcpyHostToDevice));
  if (count) NSX_HIP(ctx, hipMemcpy(ctx->d.count, count, b, hipMemcpyHostToDevice));
  if (params6) NSX_HIP(ctx, hipMemcpy(ctx->b_params, params6, b * 6, hipMemcpyHostToDevice));
  return NSX_OK;
}

nsx_status nsx_backend_upload_hh_gates(nsx_ctx* ctx, const float* m, const float* h,
                                       const float* n) {
  nsx_status s = need_backend(ctx);
  if (s) return s;
  const size_t b = ctx->n * 4;
  if (m) NSX_HIP(ctx, hipMemcpy(ctx->d.gm, m, b, hipMemcpyHostToDevice));
  if (h) NSX_HIP(ctx, hipMemcpy(ctx->d.gh, h, b, hipMemcpyHostToDevice));
  if (n) NSX_HIP(ctx, hipMemcpy(ctx->d.gn, n, b, hipMemcpyHostToDevice));
  return NSX_OK;
}

nsx_status nsx_backend_upload_csr(nsx_ctx* ctx, const uint32_t* row_ptr, const uint32_t* col,
                                  const float* weight) {
  nsx_status s = need_backend(ctx);
  if (s) return s;
  if (!row_ptr || (ctx->m && (!col || !weight)))
    return set_error(ctx, NSX_E_INVALID, "NULL CSR array");
  const uint64_t n = ctx->n;
  if (row_ptr[0] != 0 || row_ptr[n] > ctx->m)
    return set_error(ctx, NSX_E_BADCSR, "row_ptr[0] != 0 or row_ptr[n] > num_edges");
  for (uint64_t i = 0; i < n; ++i)
    if (row_ptr[i] > row_ptr[i + 1]) return set_error(ctx, NSX_E_BADCSR, "row_ptr not monotone");
  const uint64_t m = row_ptr[n];
  for (uint64_t k = 0; k < m; ++k)
    if (col[k] >= n) return set_error(ctx, NSX_E_BADCSR, "column index out of range");
  NSX_HIP(ctx, hipMemcpy(ctx->b_row_ptr32, row_ptr, (n + 1) * 4, hipMemcpyHostToDevice));
  if (m) {
    NSX_HIP(ctx, hipMemcpy(ctx->col, col, m * 4, hipMemcpyHostToDevice));
    NSX_HIP(ctx, hipMemcpy(ctx->w, weight, m * 4, hipMemcpyHostToDevice));
  }
  return NSX_OK;
}

nsx_status nsx_backend_upload_input_current(nsx_ctx* ctx, const float* input_current) {
  nsx_status s = need_backend(ctx);
  if (s) return s;
  if (!input_current) return set_error(ctx, NSX_E_INVALID, "NULL input_current");
  NSX_HIP(ctx, hipMemcpy(ctx->b_input, input_current, ctx->n * 4, hipMemcpyHostToDevice));
  return NSX_OK;
}

nsx_status nsx_backend_write_uniforms(nsx_ctx* ctx, float dt, float sim_time) {
  nsx_status s = need_backend(ctx);
  if (s) return s;
  ctx->b_dt = dt;
  ctx->b_time = sim_time;
  return NSX_OK;
}

nsx_status nsx_backend_set_model(nsx_ctx* ctx, uint32_t model) {
  nsx_status s = need_backend(ctx);
  if (s) return s;
  if (model > 2) return set_error(ctx, NSX_E_INVALID, "model must be 0 (Lif), 1 or 2");
  ctx->b_model = model;
  return NSX_OK;
}

nsx_status nsx_backend_step_neurons(nsx_ctx* ctx, uint32_t* spiked, float* v, float* u,
                                    int32_t* refr, uint32_t* count) {
  nsx_status s = need_backend(ctx);
  if (s) return s;
  BackendPtrs b = backend_ptrs(ctx);
  const uint32_t n = (uint32_t)ctx->n;
  NSX_HIP(ctx, hipMemsetAsync(ctx->d.nspk, 0, 4, ctx->stream));
  hipLaunchKernelGGL(k_backend_neurons, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, b, n,
                     ctx->b_model, ctx->b_dt, ctx->b_time);
  NSX_HIP(ctx, hipGetLastError());
  const size_t by = ctx->n * 4;
  if (spiked) NSX_HIP(ctx, hipMemcpyAsync(spiked, ctx->b_spiked, by, hipMemcpyDeviceToHost, ctx->stream));
  if (v) NSX_HIP(ctx, hipMemcpyAsync(v, ctx->d.v, by, hipMemcpyDeviceToHost, ctx->stream));
  if (u) NSX_HIP(ctx, hipMemcpyAsync(u, ctx->d.u, by, hipMemcpyDeviceToHost, ctx->stream));
  if (refr) NSX_HIP(ctx, hipMemcpyAsync(refr, ctx->d.refr, by, hipMemcpyDeviceToHost, ctx->stream));
  if (count) NSX_HIP(ctx, hipMemcpyAsync(count, ctx->d.count, by, hipMemcpyDeviceToHost, ctx->stream));
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return NSX_OK;
}

nsx_status nsx_backend_step_spmv(nsx_ctx* ctx, float* input_current) {
  nsx_status s = need_backend(ctx);
  if (s) return s;
  BackendPtrs b = backend_ptrs(ctx);
  const uint32_t n = (uint32_t)ctx->n;
  hipLaunchKernelGGL(k_backend_spmv, dim3(1024), dim3(256), 0, ctx->stream, b, ctx->p.scale_in);
  NSX_HIP(ctx, hipGetLastError());
  hipLaunchKernelGGL(k_backend_convert, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, b, n,
                     ctx->p.scale_out);
  NSX_HIP(ctx, hipGetLastError());
  if (input_current)
    NSX_HIP(ctx, hipMemcpyAsync(input_current, ctx->b_input, ctx->n * 4, hipMemcpyDeviceToHost,
                                ctx->stream));
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return NSX_OK;
}

}  // extern "C"
