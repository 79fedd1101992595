// nsx_surrogate.hip -- learned Hodgkin-Huxley surrogate (no counterpart in the reference).
//
// The surrogate replaces one network step of the plain integrator -- hh_substeps Euler sub-steps
// of dt/hh_substeps with the input current held, one evaluation of the six gate rates per
// sub-step -- by the same sub-steps with the rates frozen at the step-start V (one rate
// evaluation) plus a learned correction: an MLP 5 -> 64 -> 64 -> 4 on the normalised
// (V, m, h, n, I) predicting what the rates' change within the step adds to (V, m, h, n).
// Pieces:
//   k_surr_gen      training transitions from the plain integrator (the hh_substep shared with
//                   the network step) driven by OU currents, half with the network's drive
//                   parameters and half with random (mean, sigma, tau); the target of a state is
//                   plain step minus frozen-rate step
//   k_surr_stats    per-block moments and ranges -> normalisation and the training envelope
//   k_surr_train    forward + backward of 64-sample chunks per wavefront (bf16 MFMA, FP32
//                   accumulate), per-wave partial gradients
//   k_surr_reduce   fixed-order sum of the partials (deterministic) and the Adam update
//   k_surr_pack     FP32 masters -> bf16 per-lane MFMA fragments
//   k_update_surr   the network update: ring, conductances, drive, surrogate step, plain fallback
//                   for neurons outside the envelope, spike detection, block compaction
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "nsx_device.h"
#include "nsx_hh.h"
#include "nsx_internal.h"
#include "nsx_kernels.h"
#include "nsx_mlp.h"

namespace nsx {

constexpr uint32_t kStreamGen = 0x53555247u;    // training-data trajectories
constexpr uint32_t kStreamTrain = 0x5452414Eu;  // minibatch sampling
constexpr uint32_t kStreamInit = 0x494E4954u;   // parameter initialisation
constexpr int kSampleFloats = 9;                // V m h n I | target correction of V m h n
constexpr uint32_t kMaxTrajectories = 1u << 18;
constexpr uint32_t kStatsBlocks = 512;
constexpr int kStatQ = 28;  // 9 sums, 9 sums of squares, 5 minima, 5 maxima

struct GenParams {
  HHConst hh;
  uint32_t substeps;
  float dt;
  uint32_t T, per, warm, stride;
  uint64_t seed;
  uint32_t rand_thr;                           // trajectory takes a random drive if draw < thr
  float mu0, sig0, tau0;                       // the network's drive
  float mu_lo, mu_hi, sig_hi, tau_lo, tau_hi;  // random drives
  float scale_eps;
};

// One thread per trajectory; sample k of trajectory j is stored at (k * T + j).
__global__ __launch_bounds__(256) void k_surr_gen(GenParams g, float* data) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= g.T) return;
  const U4 r = draw(g.seed, j, 0xFFFFFFFFu, kStreamGen);
  const U4 r1 = draw(g.seed, j, 0xFFFFFFFDu, kStreamGen);
  float mu = g.mu0, sig = g.sig0, tau = g.tau0;
  if (r1.x < g.rand_thr) {
    mu = g.mu_lo + (g.mu_hi - g.mu_lo) * u01(r.x);
    sig = g.sig_hi * u01(r.y);
    tau = g.tau_lo + (g.tau_hi - g.tau_lo) * u01(r.z);
  }
  const float k1 = expf_(-g.dt / tau);
  const float k2 = mu * (1.0f - k1);
  const float k3 = sig * sqrtf(1.0f - k1 * k1);
  const U4 r0 = draw(g.seed, j, 0xFFFFFFFEu, kStreamGen);
  float I = mu + sig * normal_(r0.x, r0.y);
  float v = -65.0f + 6.0f * (u01(r.w) - 0.5f), m = 0.05f, h = 0.6f, n = 0.32f;
  const uint32_t total = g.warm + g.per * g.stride;
  for (uint32_t k = 0; k < total; ++k) {
    const float v0 = v, m0 = m, h0 = h, n0 = n;
    for (uint32_t s = 0; s < g.substeps; ++s) hh_substep(v, m, h, n, I, g.hh);
    if (k >= g.warm && (k - g.warm) % g.stride == 0) {
      float vb = v0, mb = m0, hb = h0, nb = n0;
      hh_frozen(vb, mb, hb, nb, I, g.hh, g.substeps);
      const uint32_t idx = (k - g.warm) / g.stride;
      float* o = data + ((size_t)idx * g.T + j) * kSampleFloats;
      o[0] = v0;
      o[1] = m0;
      o[2] = h0;
      o[3] = n0;
      o[4] = I;
      float sc = fabsf(vb - v0);
      sc = sc + g.scale_eps;
      o[5] = (v - vb) / sc;
      o[6] = (m - mb) / sc;
      o[7] = (h - hb) / sc;
      o[8] = (n - nb) / sc;
    }
    const U4 q = draw(g.seed, j, k, kStreamGen);
    const float z = normal_(q.x, q.y);
    I = I * k1 + k2 + k3 * z;
  }
}

// Per-block moments (double) and ranges of the samples; block b writes part[b*28 ..].
__global__ __launch_bounds__(256) void k_surr_stats(const float* data, uint64_t n,
                                                    double* part) {
  __shared__ double red[256];
  double acc[kStatQ];
  for (int q = 0; q < 18; ++q) acc[q] = 0.0;
  for (int q = 18; q < 23; ++q) acc[q] = 3.0e38;
  for (int q = 23; q < 28; ++q) acc[q] = -3.0e38;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256) {
    const float* s = data + i * kSampleFloats;
    for (int k = 0; k < kSampleFloats; ++k) {
      const double x = s[k];
      acc[k] += x;
      acc[9 + k] += x * x;
      if (k < kMlpIn) {
        acc[18 + k] = fmin(acc[18 + k], x);
        acc[23 + k] = fmax(acc[23 + k], x);
      }
    }
  }
  for (int q = 0; q < kStatQ; ++q) {
    red[threadIdx.x] = acc[q];
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
      if (threadIdx.x < w) {
        const double a = red[threadIdx.x], b = red[threadIdx.x + w];
        red[threadIdx.x] = q < 18 ? a + b : (q < 23 ? fmin(a, b) : fmax(a, b));
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)blockIdx.x * kStatQ + q] = red[0];
    __syncthreads();
  }
}

// He-uniform initialisation of the FP32 masters (biases zero); clears the Adam moments.
__global__ void k_surr_init(float* master, float* am, float* av, uint64_t seed) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (uint32_t)kMlpParams) return;
  float bound = 0.0f;
  if (p < (uint32_t)kOffB1) bound = sqrtf(6.0f / kMlpIn);
  else if (p >= (uint32_t)kOffW2 && p < (uint32_t)kOffB2) bound = sqrtf(6.0f / kMlpH);
  else if (p >= (uint32_t)kOffW3 && p < (uint32_t)kOffB3) bound = sqrtf(1.0f / kMlpH);
  const U4 r = draw(seed, p, 0u, kStreamInit);
  master[p] = bound * (2.0f * u01(r.x) - 1.0f);
  am[p] = 0.0f;
  av[p] = 0.0f;
}

// FP32 masters -> MlpImage.  Threads [0, 64 kImageFrags): one bf16x8 fragment lane each; then the
// FP32 parts (bf16-rounded layer-1 weights, biases).
__device__ inline __bf16 bf16_lo(float x) {  // the second word of a two-word bf16 split
  const float hi = (float)(__bf16)x;
  return (__bf16)(x - hi);
}

__global__ void k_surr_pack(const float* w, MlpImage* img) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < (uint32_t)kImageFrags * 64u) {
    const uint32_t f = t >> 6, l = t & 63u, half = l >> 5;
    bf16x8 r;
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) {
      __bf16 x = (__bf16)0.0f;
      if (f < 2) {  // layer 1 [h1][k]: m = h1 (tile f); k = input hi / lo column, bias words
        const uint32_t h = f * 32 + (l & 31), k = half * 8 + j;
        if (k < (uint32_t)kL1Lo) x = (__bf16)w[kOffW1 + h * kMlpIn + k];
        else if (k < (uint32_t)kL1Bias) x = (__bf16)w[kOffW1 + h * kMlpIn + k - kL1Lo];
        else if (k == (uint32_t)kL1Bias) x = (__bf16)w[kOffB1 + h];
        else if (k == (uint32_t)kL1Bias + 1) x = bf16_lo(w[kOffB1 + h]);
      } else if (f < 10) {  // W2 [h2][h1] for inference: k-step ks in accumulator row order
        const uint32_t mt = (f - 2) >> 2, ks = (f - 2) & 3u;
        x = (__bf16)w[kOffW2 + (mt * 32 + (l & 31)) * 64 + acc_row(ks, half, j)];
      } else if (f < 14) {  // [W3 hi (rows 0-3); W3 lo (rows 4-7); 0] [32][h2], k-step f - 10
        const uint32_t m = l & 31u;
        if (m < 2 * kMlpOut) {
          const float v = w[kOffW3 + (m & 3u) * 64 + acc_row(f - 10, half, j)];
          x = m < (uint32_t)kMlpOut ? (__bf16)v : bf16_lo(v);
        }
      } else if (f < 22) {  // W2 [h2][h1]: m = h2 (tile mt), k = h1 (step ks)
        const uint32_t mt = (f - 14) >> 2, ks = (f - 14) & 3u;
        x = (__bf16)w[kOffW2 + (mt * 32 + (l & 31)) * 64 + ks * 16 + half * 8 + j];
      } else if (f < 26) {  // W3 [o][h2] hi (22, 23) and lo (24, 25) words: m = o (< 4), k = h2
        const uint32_t ks = (f - 22) & 1u, o = l & 15u;
        if (o < 4) {
          const float v = w[kOffW3 + o * 64 + ks * 32 + (l >> 4) * 8 + j];
          x = f < 24 ? (__bf16)v : bf16_lo(v);
        }
      } else if (f < 34) {  // W2^T [h1][h2]: m = h1 (tile mt), k = h2 (step ks)
        const uint32_t mt = (f - 26) >> 2, ks = (f - 26) & 3u;
        const uint32_t h1 = mt * 32 + (l & 31), h2 = ks * 16 + half * 8 + j;
        x = (__bf16)w[kOffW2 + h2 * 64 + h1];
      } else {  // W3^T [h2][o] hi (34, 35) and lo (36, 37) words: m = h2, k = o (< 4)
        const uint32_t mt = (f - 34) & 1u, o = half * 8 + j;
        if (o < 4) {
          const float v = w[kOffW3 + o * 64 + mt * 32 + (l & 31)];
          x = f < 36 ? (__bf16)v : bf16_lo(v);
        }
      }
      r[j] = x;
    }
    if (f < 2) img->w1f[f][l] = r;
    else if (f < 10) img->w2pf[f - 2][l] = r;
    else if (f < 14) img->w3pf[f - 10][l] = r;
    else if (f < 22) img->w2f[f - 14][l] = r;
    else if (f < 24) img->w3f[f - 22][l] = r;
    else if (f < 26) img->w3lf[f - 24][l] = r;
    else if (f < 34) img->w2tf[f - 26][l] = r;
    else if (f < 36) img->w3tf[f - 34][l] = r;
    else img->w3tlf[f - 36][l] = r;
    return;
  }
  const uint32_t u = t - (uint32_t)kImageFrags * 64u;
  if (u < (uint32_t)kMlpH) img->b2[u] = w[kOffB2 + u];
  else if (u < (uint32_t)kScalarHead) img->b3[u - kMlpH] = w[kOffB3 + u - kMlpH];
}

struct TrainArgs {
  const MlpImage* img;
  const MlpNorm* norm;
  const float* data;
  uint64_t n;
  uint32_t cpw;    // chunks per wavefront
  uint32_t step;   // optimiser step (sampling counter)
  uint32_t fixed;  // 1: chunk c of wave w takes samples ((w*cpw + c)*64 + lane)
  uint64_t seed;
  float gscale;
  float* gpart;  // [waves][kMlpParams]
  float* lpart;  // [waves]
};

__global__ __launch_bounds__(64) void k_surr_train(TrainArgs a) {
  __shared__ TrainLds L;
  __shared__ MlpScalars S;
  const uint32_t lane = threadIdx.x, w = blockIdx.x;
  stage_scalars(S, a.img, a.norm);
  __syncthreads();
  const MlpNorm& nm = S.norm;
  GradAcc G;
  grad_zero(G);
  for (uint32_t c = 0; c < a.cpw; ++c) {
    const uint64_t gi = ((uint64_t)w * a.cpw + c) * 64 + lane;
    uint64_t idx = gi;
    if (!a.fixed) {
      const U4 r = draw(a.seed, gi, a.step, kStreamTrain);
      idx = ((uint64_t)r.x * a.n) >> 32;
    }
    const float* smp = a.data + idx * kSampleFloats;
    float xn[kMlpIn];
#pragma unroll
    for (int k = 0; k < kMlpIn; ++k) {
      float d = smp[k] - nm.in_mu[k];
      xn[k] = d * nm.in_isd[k];
      L.x[lane * kMlpIn + k] = xn[k];
    }
#pragma unroll
    for (int o = 0; o < kMlpOut; ++o) {
      float d = smp[kMlpIn + o] - nm.out_mu[o];
      L.d3[lane * 4 + o] = d / nm.out_sd[o];
    }
    wave_lds_sync();
    mlp_train_chunk(*a.img, S, L, G, xn, a.gscale, lane);
  }
  grad_store(G, a.gpart + (size_t)w * kMlpParams, lane);
  float l = G.loss;
  for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
  if (lane == 0) a.lpart[w] = l * a.gscale;
}

struct AdamArgs {
  const float* gpart;
  const float* lpart;
  uint32_t waves;
  float* master;
  float* am;
  float* av;
  float beta1, beta2, lr_eff, eps_eff;
  uint32_t adam;     // 0: write the summed gradient to gout instead of updating
  float* gout;
  float* loss_out;   // one float (the step's loss)
};

// 64 parameters per block; 16 rows of threads each sum the waves w = row (mod 16) with four
// independent accumulators (loads in flight), combined in a fixed order: deterministic.
__global__ __launch_bounds__(1024) void k_surr_reduce(AdamArgs a) {
  __shared__ float part[16][64];
  const uint32_t tx = threadIdx.x & 63u, ty = threadIdx.x >> 6;
  const uint32_t p = blockIdx.x * 64 + tx;
  float g0 = 0.0f, g1 = 0.0f, g2 = 0.0f, g3 = 0.0f;
  if (p < (uint32_t)kMlpParams) {
    const float* src = a.gpart + p;
    uint32_t w = ty;
    for (; w + 48 < a.waves; w += 64) {
      g0 += src[(size_t)w * kMlpParams];
      g1 += src[(size_t)(w + 16) * kMlpParams];
      g2 += src[(size_t)(w + 32) * kMlpParams];
      g3 += src[(size_t)(w + 48) * kMlpParams];
    }
    for (; w < a.waves; w += 16) g0 += src[(size_t)w * kMlpParams];
  }
  part[ty][tx] = (g0 + g1) + (g2 + g3);
  __syncthreads();
  if (ty == 0 && p < (uint32_t)kMlpParams) {
    float g = 0.0f;
#pragma unroll
    for (int r = 0; r < 16; ++r) g += part[r][tx];
    if (a.adam) {
      const float m = a.beta1 * a.am[p] + (1.0f - a.beta1) * g;
      const float v = a.beta2 * a.av[p] + (1.0f - a.beta2) * g * g;
      a.am[p] = m;
      a.av[p] = v;
      a.master[p] -= a.lr_eff * m / (sqrtf(v) + a.eps_eff);
    } else {
      a.gout[p] = g;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && a.loss_out) {
    float l = 0.0f;
    for (uint32_t w = 0; w < a.waves; ++w) l += a.lpart[w];
    *a.loss_out = l;
  }
}

// One surrogate step of a neuron with raw state x = (V, m, h, n, I): next (V, m, h, n) in nx, the
// MLP's output in target units in res (the correction is res * (|dV| + scale_eps), dV the
// frozen-rate step's change of V), and whether x lay inside the training envelope
// (otherwise nx is the plain step).  Every lane of the wavefront must call it; `live` lanes only
// use the result.
template <class WT>
__device__ inline bool surr_step(const WT& W, const MlpScalars& S, const HHConst& c,
                                 uint32_t substeps, const float x[kMlpIn], bool live,
                                 float nx[4], float res[kMlpOut], uint32_t lane) {
  const MlpNorm& nm = S.norm;
  bool in = live;
  float xn[kMlpIn];
#pragma unroll
  for (int k = 0; k < kMlpIn; ++k) {
    in = in && x[k] >= nm.lo[k] && x[k] <= nm.hi[k];
    float d = x[k] - nm.in_mu[k];
    xn[k] = d * nm.in_isd[k];
  }
  float y[kMlpOut];
  mlp_infer(W, S, xn, y, lane);  // the MFMAs span the wavefront
#pragma unroll
  for (int o = 0; o < kMlpOut; ++o) {
    float d = y[o] * nm.out_sd[o];
    res[o] = d + nm.out_mu[o];
  }
  float v = x[0], m = x[1], h = x[2], n = x[3];
  if (in) {
    hh_frozen(v, m, h, n, x[4], c, substeps);
    float sc = fabsf(v - x[0]);
    sc = sc + nm.scale_eps;
    v = v + res[0] * sc;
    m = clampf_(m + res[1] * sc, 0.0f, 1.0f);
    h = clampf_(h + res[2] * sc, 0.0f, 1.0f);
    n = clampf_(n + res[3] * sc, 0.0f, 1.0f);
  } else if (live) {  // outside the training envelope: the plain step
    for (uint32_t k = 0; k < substeps; ++k) hh_substep(v, m, h, n, x[4], c);
  }
  nx[0] = v;
  nx[1] = m;
  nx[2] = h;
  nx[3] = n;
  return in;
}

// Surrogate on caller-supplied raw states (count x 5).  what: 0 the MLP's output (target units),
// 1 the surrogate's next state (V, m, h, n) as the network step computes it, 2 the next state of
// the frozen-rate sub-steps alone.  inside: envelope flags.
__global__ __launch_bounds__(256) void k_surr_forward(const MlpImage* img, const MlpNorm* norm,
                                                      HHConst c, uint32_t substeps,
                                                      const float* x, uint64_t count,
                                                      uint32_t what, float* y, uint8_t* inside) {
  __shared__ MlpScalars S;
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  const uint32_t lane = threadIdx.x & 63u;
  stage_scalars(S, img, norm);
  __syncthreads();
  const MlpNorm& nm = S.norm;
  const bool live = i < count;
  float xs[kMlpIn];
#pragma unroll
  for (int k = 0; k < kMlpIn; ++k) xs[k] = live ? x[i * kMlpIn + k] : nm.in_mu[k];
  float nx[4], res[kMlpOut];
  const bool in = surr_step(*img, S, c, substeps, xs, live, nx, res, lane);
  if (live) {
    if (what == 2) {
      nx[0] = xs[0];
      nx[1] = xs[1];
      nx[2] = xs[2];
      nx[3] = xs[3];
      hh_frozen(nx[0], nx[1], nx[2], nx[3], xs[4], c, substeps);
    }
#pragma unroll
    for (int o = 0; o < 4; ++o) y[i * 4 + o] = what == 0 ? res[o] : nx[o];
    if (inside) inside[i] = in ? 1 : 0;
  }
}

struct SurrDev {
  const MlpImage* img;
  const MlpNorm* norm;
  unsigned long long* fb;  // per-block fallback counts
};

// Network update of one step with the surrogate (HH conductance model, crossing spikes).
template <int DRIVE, bool PAIR>
__global__ __launch_bounds__(kBlock) void k_update_surr(StepParams p, DevPtrs s, SurrDev q,
                                                        int32_t off) {
  __shared__ MlpScalars S;
  __shared__ MlpInferW WL;
  __shared__ uint32_t s_fb[kBlock / 64];
  __shared__ uint32_t s_drive[kBlock];
  const int32_t t = *s.tbase + off;
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
  const NeuronId id = neuron_id(p, s, i, blockIdx.x);
  const bool live = i < p.n && id.local < id.n;
  // the neuron's loads first, in flight together while the block stages its scalars and drive
  RingIn rg{nullptr, 0, 0};
  HHSyn y{0.0f, 0.0f, 0.0f, 0.0f};
  float x[kMlpIn];
  int32_t last = kNeverStep;
  const uint32_t prev = PAIR ? s.bcnt[blockIdx.x] : 0u;  // 0: no spike at t-1 in the block
  const bool ring_in = ring_active(p, s, t);
  if (live) {
    rg = ring_load<true>(p, s, i, t, ring_in);
    if (PAIR && prev) last = s.last[i];
    x[0] = s.v[i];
    x[1] = s.gm[i];
    x[2] = s.gh[i];
    x[3] = s.gn[i];
    y = hh_syn_load<DRIVE>(s, i);
  }
  stage_scalars(S, q.img, q.norm);
  stage_infer_weights(WL, q.img);  // one L2 read per block instead of one per wavefront
  drive_stage<DRIVE>(p, s, t, s_drive, id);
  __syncthreads();
  const MlpNorm& nm = S.norm;
  float lfp = 0.0f;
  if (live) {
    post_bump_prev<PAIR>(p, s, i, last, t);
    float a0, a1;
    ring_take<true>(p, rg, a0, a1);
    float z = 0.0f;
    drive_draw<DRIVE>(p, s_drive, a0, z);
    x[4] = hh_input<DRIVE>(p, y, x[0], a0, a1, z, lfp);
    hh_syn_store<DRIVE>(s, i, y);
  } else {
#pragma unroll
    for (int k = 0; k < kMlpIn; ++k) x[k] = nm.in_mu[k];
  }
  float nx[4], res[kMlpOut];
  const bool in = surr_step(WL, S, p.hh, p.hh_substeps, x, live, nx, res, lane);
  bool spiked = false;
  if (live) {
    spiked = (x[0] < p.hh_spike_v) && (nx[0] >= p.hh_spike_v);
    s.v[i] = nx[0];
    s.gm[i] = nx[1];
    s.gh[i] = nx[2];
    s.gn[i] = nx[3];
    if (spiked) spike_record<PAIR ? NSX_PLAST_PAIR : NSX_PLAST_NONE>(p, s, i, t);
  }
  if (p.stats_on) lfp_accumulate(s, lfp, blockIdx.x);
  const unsigned long long fbm = __ballot(live && !in);
  if (lane == 0) s_fb[wid] = (uint32_t)__popcll(fbm);
  block_compact(p, s, spiked, i, t, blockIdx.x);  // synchronises the block
  if (threadIdx.x == 0) {
    uint32_t cnt = 0;
#pragma unroll
    for (uint32_t k = 0; k < kBlock / 64; ++k) cnt += s_fb[k];
    if (cnt) q.fb[blockIdx.x] += cnt;
  }
}

typedef void (*SurrUpdateFn)(StepParams, DevPtrs, SurrDev, int32_t);

template <int DRIVE>
static SurrUpdateFn surr_pick_pair(bool pair) {
  return pair ? k_update_surr<DRIVE, true> : k_update_surr<DRIVE, false>;
}

hipError_t launch_update_surrogate(nsx_ctx* ctx, int32_t off) {
  const bool pair = ctx->cfg.plasticity == NSX_PLAST_PAIR;
  SurrUpdateFn f;
  switch (ctx->cfg.drive) {
    case NSX_DRIVE_POISSON: f = surr_pick_pair<NSX_DRIVE_POISSON>(pair); break;
    case NSX_DRIVE_GAUSSIAN: f = surr_pick_pair<NSX_DRIVE_GAUSSIAN>(pair); break;
    case NSX_DRIVE_OU: f = surr_pick_pair<NSX_DRIVE_OU>(pair); break;
    default: f = surr_pick_pair<NSX_DRIVE_NONE>(pair); break;
  }
  const SurrDev q{(const MlpImage*)ctx->sur.image, (const MlpNorm*)ctx->sur.norm, ctx->sur.fb};
  const uint32_t gu = (uint32_t)((ctx->n + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(f, dim3(gu), dim3(kBlock), 0, ctx->stream, ctx->p, ctx->d, q, off);
  return hipGetLastError();
}

// Epoch stepping (networks without plasticity, shortest delay d_min >= 2 steps): `len` steps of
// k_update_surr in one launch, each neuron's (V, m, h, n), conductances and OU value in registers,
// the MLP's scalars staged once, each step's spikes into its own segment (k_synapses_epoch
// propagates them).  Every step is k_update_surr's arithmetic, so the run is bit-identical.
// No waves-per-EU bound: at its natural ~160 VGPRs (three wavefronts per SIMD, no VGPR spills)
// the kernel runs 11.3 us per step on 200k neurons against 11.6 us when held to four wavefronts
// with spills.
template <int DRIVE>
__global__ __launch_bounds__(kBlock) void k_epoch_surr(StepParams p, DevPtrs s, SurrDev q,
                                                       EpochSeg ep, int32_t off, uint32_t len) {
  __shared__ MlpScalars S;
  __shared__ MlpInferW WL;
  __shared__ uint32_t s_fb[kBlock / 64];
  __shared__ uint32_t s_drive[kBlock];
  const int32_t t0 = *s.tbase + off;
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
  const NeuronId id = neuron_id(p, s, i, blockIdx.x);
  const bool live = i < p.n && id.local < id.n;
  int32_t m0 = kNeverStep;
#pragma unroll
  for (uint32_t k = 0; k < kEvStripes; ++k) m0 = max(m0, s.ev_last[k * kEvStride]);
  HHSyn y{0.0f, 0.0f, 0.0f, 0.0f};
  float st[4] = {0.0f, 0.0f, 0.0f, 0.0f};  // V, m, h, n
  if (live) {
    st[0] = s.v[i];
    st[1] = s.gm[i];
    st[2] = s.gh[i];
    st[3] = s.gn[i];
    y = hh_syn_load<DRIVE>(s, i);
  }
  stage_scalars(S, q.img, q.norm);
  stage_infer_weights(WL, q.img);
  const size_t plane = (size_t)p.ch * p.n;
  long long qa = 0, qb = 0;
  if (live && m0 >= t0 - (int32_t)p.D + 1) {
    const long long* r = s.ring + (size_t)((uint32_t)t0 % p.D) * plane + i;
    qa = r[0];
    qb = r[p.n];
  }
  uint32_t nspk = 0, wave_fb = 0;
  int32_t last = kNeverStep;
  for (uint32_t k = 0; k < len; ++k) {
    const int32_t t = t0 + (int32_t)k;
    long long* rk = s.ring + (size_t)((uint32_t)t % p.D) * plane + i;
    const long long a = qa, b = qb;
    qa = 0;
    qb = 0;
    if (live && k + 1 < len && m0 >= t + 1 - (int32_t)p.D + 1) {
      const long long* rn = s.ring + (size_t)((uint32_t)(t + 1) % p.D) * plane + i;
      qa = rn[0];
      qb = rn[p.n];
    }
    // keeps the next step's ring loads issued here, a whole step ahead of their use
    asm volatile("" ::: "memory");
    __syncthreads();  // the previous step's readers of the block's LDS are done
    drive_stage<DRIVE>(p, s, t, s_drive, id);
    __syncthreads();
    const MlpNorm& nm = S.norm;
    float lfp = 0.0f;
    float x[kMlpIn];
    if (live) {
      float a0 = dequantize(a, p.scale_out);
      if (a != 0) rk[0] = 0;
      float a1 = dequantize(b, p.scale_out);
      if (b != 0) rk[p.n] = 0;
      float z = 0.0f;
      drive_draw<DRIVE>(p, s_drive, a0, z);
      x[4] = hh_input<DRIVE>(p, y, st[0], a0, a1, z, lfp);
      x[0] = st[0];
      x[1] = st[1];
      x[2] = st[2];
      x[3] = st[3];
    } else {
#pragma unroll
      for (int kk = 0; kk < kMlpIn; ++kk) x[kk] = nm.in_mu[kk];
    }
    float nx[4], res[kMlpOut];
    const bool in = surr_step(WL, S, p.hh, p.hh_substeps, x, live, nx, res, lane);
    bool spiked = false;
    if (live) {
      spiked = (x[0] < p.hh_spike_v) && (nx[0] >= p.hh_spike_v);
      st[0] = nx[0];
      st[1] = nx[1];
      st[2] = nx[2];
      st[3] = nx[3];
      if (spiked) {
        nspk += 1u;
        last = t;
      }
    }
    if (p.stats_on) lfp_accumulate(s, lfp, blockIdx.x);
    wave_fb += (uint32_t)__popcll(__ballot(live && !in));
    DevPtrs sk = s;
    sk.spikes = ep.spikes + ((size_t)k * p.nblk) * kBlock;
    sk.bcnt = ep.bcnt + (size_t)k * p.nblk;
    block_compact(p, sk, spiked, i, t, blockIdx.x);
  }
  if (live) {
    s.v[i] = st[0];
    s.gm[i] = st[1];
    s.gh[i] = st[2];
    s.gn[i] = st[3];
    hh_syn_store<DRIVE>(s, i, y);
    if (nspk) {
      s.count[i] += nspk;
      s.last[i] = last;
    }
  }
  if (lane == 0) s_fb[wid] = wave_fb;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t cnt = 0;
#pragma unroll
    for (uint32_t k = 0; k < kBlock / 64; ++k) cnt += s_fb[k];
    if (cnt) q.fb[blockIdx.x] += cnt;
  }
}

typedef void (*SurrEpochFn)(StepParams, DevPtrs, SurrDev, EpochSeg, int32_t, uint32_t);

hipError_t launch_epoch_surrogate(nsx_ctx* ctx, int32_t off, uint32_t len) {
  SurrEpochFn f;
  switch (ctx->cfg.drive) {
    case NSX_DRIVE_POISSON: f = k_epoch_surr<NSX_DRIVE_POISSON>; break;
    case NSX_DRIVE_GAUSSIAN: f = k_epoch_surr<NSX_DRIVE_GAUSSIAN>; break;
    case NSX_DRIVE_OU: f = k_epoch_surr<NSX_DRIVE_OU>; break;
    default: f = k_epoch_surr<NSX_DRIVE_NONE>; break;
  }
  const SurrDev q{(const MlpImage*)ctx->sur.image, (const MlpNorm*)ctx->sur.norm, ctx->sur.fb};
  hipLaunchKernelGGL(f, dim3(ctx->p.nblk), dim3(kBlock), 0, ctx->stream, ctx->p, ctx->d, q,
                     ctx->ep, off, len);
  return hipGetLastError();
}

// ---- host side ------------------------------------------------------------------------------

static nsx_status need_hh(nsx_ctx* ctx) {
  if (!ctx) return set_error(nullptr, NSX_E_INVALID, "ctx is NULL");
  if (ctx->cfg.model != NSX_MODEL_HH && ctx->cfg.model != NSX_MODEL_HH_SURROGATE)
    return set_error(ctx, NSX_E_INVALID, "the surrogate needs a Hodgkin-Huxley context");
  if (ctx->p.hh.literal)
    return set_error(ctx, NSX_E_INVALID,
                     "the surrogate models the documented HH equation (crossing spike mode)");
  return NSX_OK;
}

static nsx_status sur_alloc(nsx_ctx* ctx) {
  Surrogate& S = ctx->sur;
  nsx_status s;
  if (!S.master) {
    if ((s = dev_alloc(ctx, (void**)&S.master, kMlpParams * 4, true))) return s;
    if ((s = dev_alloc(ctx, (void**)&S.adam_m, kMlpParams * 4, true))) return s;
    if ((s = dev_alloc(ctx, (void**)&S.adam_v, kMlpParams * 4, true))) return s;
    if ((s = dev_alloc(ctx, &S.image, sizeof(MlpImage), true))) return s;
    if ((s = dev_alloc(ctx, (void**)&S.norm, sizeof(MlpNorm), true))) return s;
    if ((s = dev_alloc(ctx, (void**)&S.loss, 4, true))) return s;
  }
  return NSX_OK;
}

static nsx_status sur_parts(nsx_ctx* ctx, uint32_t waves) {
  Surrogate& S = ctx->sur;
  if (S.gpart_waves >= waves) return NSX_OK;
  if (S.gpart) dev_free(ctx, S.gpart);
  if (S.lpart) dev_free(ctx, S.lpart);
  S.gpart = S.lpart = nullptr;
  S.gpart_waves = 0;
  nsx_status s;
  if ((s = dev_alloc(ctx, (void**)&S.gpart, (size_t)waves * kMlpParams * 4, false))) return s;
  if ((s = dev_alloc(ctx, (void**)&S.lpart, (size_t)waves * 4, false))) return s;
  S.gpart_waves = waves;
  return NSX_OK;
}

static hipError_t sur_pack(nsx_ctx* ctx) {
  hipLaunchKernelGGL(k_surr_pack, dim3((kImageFrags * 64 + kScalarHead + 255) / 256), dim3(256), 0,
                     ctx->stream,
                     ctx->sur.master, (MlpImage*)ctx->sur.image);
  return hipGetLastError();
}

}  // namespace nsx

using namespace nsx;

extern "C" {

nsx_status nsx_surrogate_generate(nsx_ctx* ctx, const nsx_gen_config* gc, uint64_t* n_out) {
  nsx_status s = need_hh(ctx);
  if (s) return s;
  if (!gc || gc->struct_size != sizeof(nsx_gen_config))
    return set_error(ctx, NSX_E_INVALID, "nsx_gen_config missing or struct_size mismatch");
  const uint64_t n_samples = gc->n_samples;
  if (n_samples < 64) return set_error(ctx, NSX_E_INVALID, "n_samples must be >= 64");
  if (gc->stride < 1 || !(gc->warm_ms >= 0) || !(gc->scale_eps > 0) ||
      !(gc->random_fraction >= 0 && gc->random_fraction <= 1) || !(gc->mu_lo <= gc->mu_hi) ||
      !(gc->sigma_hi >= 0) || !(gc->tau_lo > 0 && gc->tau_lo <= gc->tau_hi))
    return set_error(ctx, NSX_E_INVALID, "invalid generator parameters");
  NSX_HIP(ctx, hipSetDevice(ctx->device));
  if ((s = sur_alloc(ctx))) return s;
  Surrogate& S = ctx->sur;
  uint32_t T = 1;
  while (T * 2 <= kMaxTrajectories && (uint64_t)T * 2 <= n_samples) T *= 2;
  const uint64_t per = n_samples / T;
  if (per > 0xffffffffull / 64 || per * gc->stride > 0x7fffffffull)
    return set_error(ctx, NSX_E_INVALID, "n_samples * stride too large");
  const uint64_t N = per * T;
  if (S.data) dev_free(ctx, S.data);
  S.data = nullptr;
  S.n_data = 0;
  if ((s = dev_alloc(ctx, (void**)&S.data, (size_t)N * kSampleFloats * 4, false))) return s;
  const nsx_config& c = ctx->cfg;
  GenParams g{};
  g.hh = ctx->p.hh;
  g.substeps = ctx->p.hh_substeps;
  g.dt = (float)c.dt_ms;
  g.T = T;
  g.per = (uint32_t)per;
  g.warm = (uint32_t)ceil(gc->warm_ms / c.dt_ms);
  g.stride = gc->stride;
  g.seed = gc->seed;
  const double thr = floor((double)gc->random_fraction * 4294967296.0);
  g.rand_thr = thr >= 4294967295.0 ? 0xffffffffu : (uint32_t)thr;
  const bool ou = c.drive == NSX_DRIVE_OU;
  g.mu0 = ou ? c.drive_mean : 6.0f;
  g.sig0 = ou ? c.drive_sigma : 3.0f;
  g.tau0 = ou ? c.drive_tau_ms : 5.0f;
  g.mu_lo = gc->mu_lo;
  g.mu_hi = gc->mu_hi;
  g.sig_hi = gc->sigma_hi;
  g.tau_lo = gc->tau_lo;
  g.tau_hi = gc->tau_hi;
  g.scale_eps = gc->scale_eps;
  hipLaunchKernelGGL(k_surr_gen, dim3((T + 255) / 256), dim3(256), 0, ctx->stream, g, S.data);
  NSX_HIP(ctx, hipGetLastError());
  // normalisation and envelope
  double* part = nullptr;
  if ((s = dev_alloc(ctx, (void**)&part, (size_t)kStatsBlocks * kStatQ * 8, false))) return s;
  hipLaunchKernelGGL(k_surr_stats, dim3(kStatsBlocks), dim3(256), 0, ctx->stream, S.data, N,
                     part);
  hipError_t e = hipGetLastError();
  std::vector<double> h((size_t)kStatsBlocks * kStatQ);
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), part, h.size() * 8, hipMemcpyDeviceToHost,
                                          ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  dev_free(ctx, part);
  if (e != hipSuccess) return hip_error(ctx, e, "surrogate data statistics");
  double tot[kStatQ];
  for (int q = 0; q < kStatQ; ++q) tot[q] = q < 18 ? 0.0 : (q < 23 ? 3.0e38 : -3.0e38);
  for (uint32_t b = 0; b < kStatsBlocks; ++b)
    for (int q = 0; q < kStatQ; ++q) {
      const double x = h[(size_t)b * kStatQ + q];
      tot[q] = q < 18 ? tot[q] + x : (q < 23 ? std::min(tot[q], x) : std::max(tot[q], x));
    }
  MlpNorm nm{};
  nm.scale_eps = gc->scale_eps;
  for (int k = 0; k < kSampleFloats; ++k) {
    const double mean = tot[k] / (double)N;
    const double var = std::max(tot[9 + k] / (double)N - mean * mean, 0.0);
    const double sd = var > 1e-24 ? sqrt(var) : 1.0;
    if (k < kMlpIn) {
      nm.in_mu[k] = (float)mean;
      nm.in_isd[k] = (float)(1.0 / sd);
      nm.lo[k] = (float)tot[18 + k];
      nm.hi[k] = (float)tot[23 + k];
    } else {
      nm.out_mu[k - kMlpIn] = (float)mean;
      nm.out_sd[k - kMlpIn] = (float)sd;
    }
  }
  if (!std::isfinite(tot[0]) || !std::isfinite(tot[9]))
    return set_error(ctx, NSX_E_STATE, "training transitions are not finite");
  NSX_HIP(ctx, hipMemcpy(S.norm, &nm, sizeof(nm), hipMemcpyHostToDevice));
  S.n_data = N;
  S.has_norm = true;
  if (n_out) *n_out = N;
  return NSX_OK;
}

nsx_status nsx_surrogate_train(nsx_ctx* ctx, const nsx_train_config* tc, float* loss_history) {
  nsx_status s = need_hh(ctx);
  if (s) return s;
  if (!tc || tc->struct_size != sizeof(nsx_train_config))
    return set_error(ctx, NSX_E_INVALID, "nsx_train_config missing or struct_size mismatch");
  Surrogate& S = ctx->sur;
  if (!S.data || !S.has_norm)
    return set_error(ctx, NSX_E_STATE, "no training data (call nsx_surrogate_generate)");
  if (tc->waves < 1 || tc->waves > 65536 || tc->chunks_per_wave < 1 || tc->steps < 1)
    return set_error(ctx, NSX_E_INVALID, "waves, chunks_per_wave and steps must be >= 1");
  if (!(tc->lr > 0) || !(tc->beta1 >= 0 && tc->beta1 < 1) || !(tc->beta2 >= 0 && tc->beta2 < 1) ||
      !(tc->eps > 0))
    return set_error(ctx, NSX_E_INVALID, "invalid optimiser constants");
  if (S.n_data >= (1ull << 32))
    return set_error(ctx, NSX_E_INVALID, "training set above 2^32 samples");
  NSX_HIP(ctx, hipSetDevice(ctx->device));
  if ((s = sur_alloc(ctx))) return s;
  if ((s = sur_parts(ctx, tc->waves))) return s;
  float* hist = nullptr;
  if ((s = dev_alloc(ctx, (void**)&hist, (size_t)tc->steps * 4, true))) return s;
  if (tc->init || !S.has_params) {
    hipLaunchKernelGGL(k_surr_init, dim3((kMlpParams + 255) / 256), dim3(256), 0, ctx->stream,
                       S.master, S.adam_m, S.adam_v, tc->seed);
    S.adam_t = 0;
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = sur_pack(ctx);
  TrainArgs ta{};
  ta.img = (const MlpImage*)S.image;
  ta.norm = (const MlpNorm*)S.norm;
  ta.data = S.data;
  ta.n = S.n_data;
  ta.cpw = tc->chunks_per_wave;
  ta.fixed = 0;
  ta.seed = tc->seed;
  ta.gscale = (float)(1.0 / ((double)tc->waves * tc->chunks_per_wave * 64.0 * kMlpOut));
  ta.gpart = S.gpart;
  ta.lpart = S.lpart;
  AdamArgs aa{};
  aa.gpart = S.gpart;
  aa.lpart = S.lpart;
  aa.waves = tc->waves;
  aa.master = S.master;
  aa.am = S.adam_m;
  aa.av = S.adam_v;
  aa.beta1 = tc->beta1;
  aa.beta2 = tc->beta2;
  aa.adam = 1;
  const double lr_min = tc->lr_min > 0 ? tc->lr_min : 0.0;
  for (uint32_t k = 0; k < tc->steps && e == hipSuccess; ++k) {
    const uint64_t at = ++S.adam_t;
    const double lr = lr_min + 0.5 * ((double)tc->lr - lr_min) *
                                   (1.0 + cos(M_PI * (double)k / (double)tc->steps));
    const double c1 = 1.0 - pow((double)tc->beta1, (double)at);
    const double c2 = 1.0 - pow((double)tc->beta2, (double)at);
    aa.lr_eff = (float)(lr * sqrt(c2) / c1);
    aa.eps_eff = (float)(tc->eps * sqrt(c2));
    aa.loss_out = hist + k;
    ta.step = (uint32_t)(at - 1) + tc->sample_offset;
    hipLaunchKernelGGL(k_surr_train, dim3(tc->waves), dim3(64), 0, ctx->stream, ta);
    hipLaunchKernelGGL(k_surr_reduce, dim3((kMlpParams + 63) / 64), dim3(1024), 0, ctx->stream,
                       aa);
    e = hipGetLastError();
    if (e == hipSuccess) e = sur_pack(ctx);
  }
  if (e == hipSuccess && loss_history)
    e = hipMemcpyAsync(loss_history, hist, (size_t)tc->steps * 4, hipMemcpyDeviceToHost,
                       ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  dev_free(ctx, hist);
  if (e != hipSuccess) return hip_error(ctx, e, "surrogate training");
  S.has_params = true;
  return NSX_OK;
}

nsx_status nsx_surrogate_get(nsx_ctx* ctx, float* params, float* norm) {
  nsx_status s = need_hh(ctx);
  if (s) return s;
  Surrogate& S = ctx->sur;
  if (params && !S.has_params) return set_error(ctx, NSX_E_STATE, "no surrogate parameters");
  if (norm && !S.has_norm) return set_error(ctx, NSX_E_STATE, "no surrogate normalisation");
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (params) NSX_HIP(ctx, hipMemcpy(params, S.master, kMlpParams * 4, hipMemcpyDeviceToHost));
  if (norm) NSX_HIP(ctx, hipMemcpy(norm, S.norm, sizeof(MlpNorm), hipMemcpyDeviceToHost));
  return NSX_OK;
}

nsx_status nsx_surrogate_set(nsx_ctx* ctx, const float* params, const float* norm) {
  nsx_status s = need_hh(ctx);
  if (s) return s;
  if (!params || !norm) return set_error(ctx, NSX_E_INVALID, "params and norm are required");
  const MlpNorm* nm = reinterpret_cast<const MlpNorm*>(norm);
  for (int k = 0; k < kMlpOut; ++k)
    if (!(nm->out_sd[k] > 0.0f)) return set_error(ctx, NSX_E_INVALID, "out_sd must be > 0");
  for (int k = 0; k < kMlpIn; ++k)
    if (!(nm->lo[k] <= nm->hi[k])) return set_error(ctx, NSX_E_INVALID, "envelope lo > hi");
  if (!(nm->scale_eps > 0.0f)) return set_error(ctx, NSX_E_INVALID, "scale_eps must be > 0");
  NSX_HIP(ctx, hipSetDevice(ctx->device));
  if ((s = sur_alloc(ctx))) return s;
  Surrogate& S = ctx->sur;
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  NSX_HIP(ctx, hipMemcpy(S.master, params, kMlpParams * 4, hipMemcpyHostToDevice));
  NSX_HIP(ctx, hipMemcpy(S.norm, norm, sizeof(MlpNorm), hipMemcpyHostToDevice));
  NSX_HIP(ctx, hipMemset(S.adam_m, 0, kMlpParams * 4));
  NSX_HIP(ctx, hipMemset(S.adam_v, 0, kMlpParams * 4));
  S.adam_t = 0;
  NSX_HIP(ctx, sur_pack(ctx));
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  S.has_params = S.has_norm = true;
  return NSX_OK;
}

nsx_status nsx_surrogate_forward(nsx_ctx* ctx, const float* x, uint64_t count, uint32_t what,
                                 float* y, uint8_t* inside) {
  nsx_status s = need_hh(ctx);
  if (s) return s;
  Surrogate& S = ctx->sur;
  if (!S.has_params || !S.has_norm) return set_error(ctx, NSX_E_STATE, "no surrogate parameters");
  if (count == 0) return NSX_OK;
  if (!x || !y) return set_error(ctx, NSX_E_INVALID, "x and y are required");
  if (what > 2) return set_error(ctx, NSX_E_INVALID, "what must be 0, 1 or 2");
  NSX_HIP(ctx, hipSetDevice(ctx->device));
  float *dx = nullptr, *dd = nullptr;
  uint8_t* di = nullptr;
  if ((s = dev_alloc(ctx, (void**)&dx, count * kMlpIn * 4, false))) return s;
  if ((s = dev_alloc(ctx, (void**)&dd, count * kMlpOut * 4, false))) {
    dev_free(ctx, dx);
    return s;
  }
  if ((s = dev_alloc(ctx, (void**)&di, count, false))) {
    dev_free(ctx, dx);
    dev_free(ctx, dd);
    return s;
  }
  hipError_t e = hipMemcpyAsync(dx, x, count * kMlpIn * 4, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_surr_forward, dim3((uint32_t)((count + 255) / 256)), dim3(256), 0,
                       ctx->stream, (const MlpImage*)S.image, (const MlpNorm*)S.norm, ctx->p.hh,
                       ctx->p.hh_substeps, dx, count, what, dd, di);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(y, dd, count * kMlpOut * 4, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess && inside)
    e = hipMemcpyAsync(inside, di, count, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  dev_free(ctx, dx);
  dev_free(ctx, dd);
  dev_free(ctx, di);
  if (e != hipSuccess) return hip_error(ctx, e, "surrogate forward");
  return NSX_OK;
}

nsx_status nsx_surrogate_gradient(nsx_ctx* ctx, const float* x, const float* target,
                                  uint64_t count, float* grad, float* loss) {
  nsx_status s = need_hh(ctx);
  if (s) return s;
  Surrogate& S = ctx->sur;
  if (!S.has_params || !S.has_norm) return set_error(ctx, NSX_E_STATE, "no surrogate parameters");
  if (count == 0 || count % 64 != 0 || count > (64ull << 16))
    return set_error(ctx, NSX_E_INVALID, "count must be a positive multiple of 64, <= 2^22");
  if (!x || !target || !grad) return set_error(ctx, NSX_E_INVALID, "x, target, grad required");
  NSX_HIP(ctx, hipSetDevice(ctx->device));
  const uint32_t waves = (uint32_t)(count / 64);
  if ((s = sur_parts(ctx, waves))) return s;
  std::vector<float> h((size_t)count * kSampleFloats);
  for (uint64_t i = 0; i < count; ++i) {
    memcpy(&h[i * kSampleFloats], x + i * kMlpIn, kMlpIn * 4);
    memcpy(&h[i * kSampleFloats + kMlpIn], target + i * kMlpOut, kMlpOut * 4);
  }
  float *dd = nullptr, *dg = nullptr;
  if ((s = dev_alloc(ctx, (void**)&dd, h.size() * 4, false))) return s;
  if ((s = dev_alloc(ctx, (void**)&dg, (kMlpParams + 1) * 4, false))) {
    dev_free(ctx, dd);
    return s;
  }
  hipError_t e = hipMemcpyAsync(dd, h.data(), h.size() * 4, hipMemcpyHostToDevice, ctx->stream);
  TrainArgs ta{};
  ta.img = (const MlpImage*)S.image;
  ta.norm = (const MlpNorm*)S.norm;
  ta.data = dd;
  ta.n = count;
  ta.cpw = 1;
  ta.fixed = 1;
  ta.gscale = (float)(1.0 / ((double)count * kMlpOut));
  ta.gpart = S.gpart;
  ta.lpart = S.lpart;
  AdamArgs aa{};
  aa.gpart = S.gpart;
  aa.lpart = S.lpart;
  aa.waves = waves;
  aa.adam = 0;
  aa.gout = dg;
  aa.loss_out = dg + kMlpParams;
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_surr_train, dim3(waves), dim3(64), 0, ctx->stream, ta);
    hipLaunchKernelGGL(k_surr_reduce, dim3((kMlpParams + 63) / 64), dim3(1024), 0, ctx->stream,
                       aa);
    e = hipGetLastError();
  }
  float lh = 0.0f;
  if (e == hipSuccess)
    e = hipMemcpyAsync(grad, dg, kMlpParams * 4, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(&lh, dg + kMlpParams, 4, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  dev_free(ctx, dd);
  dev_free(ctx, dg);
  if (e != hipSuccess) return hip_error(ctx, e, "surrogate gradient");
  if (loss) *loss = lh;
  return NSX_OK;
}

nsx_status nsx_surrogate_read_data(nsx_ctx* ctx, uint64_t first, uint64_t count, float* out) {
  nsx_status s = need_hh(ctx);
  if (s) return s;
  Surrogate& S = ctx->sur;
  if (first > S.n_data || count > S.n_data - first)
    return set_error(ctx, NSX_E_INVALID, "sample range outside the training set");
  if (count == 0) return NSX_OK;
  if (!out) return set_error(ctx, NSX_E_INVALID, "out is NULL");
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  NSX_HIP(ctx, hipMemcpy(out, S.data + first * kSampleFloats, count * kSampleFloats * 4,
                         hipMemcpyDeviceToHost));
  return NSX_OK;
}

nsx_status nsx_surrogate_fallbacks(nsx_ctx* ctx, uint64_t* count) {
  if (!ctx || !count) return set_error(ctx, NSX_E_INVALID, "ctx and count are required");
  *count = 0;
  if (!ctx->sur.fb) return NSX_OK;
  NSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<unsigned long long> h(ctx->p.nblk);
  NSX_HIP(ctx, hipMemcpy(h.data(), ctx->sur.fb, h.size() * 8, hipMemcpyDeviceToHost));
  uint64_t c = 0;
  for (unsigned long long x : h) c += x;
  *count = c;
  return NSX_OK;
}

}  // extern "C"
