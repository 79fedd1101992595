/*
 * oracle.c -- CPU oracle for the NeuralSim simulation core.  TEST INFRASTRUCTURE ONLY
 * (checker for tests/, __graft_entry__.smoke() and bench.py's cpu_baseline); the product
 * never links it.  Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no -ffast-math, so
 * float expressions round exactly as written, one IEEE operation at a time).
 *
 * Citations are to the reference tree (Thenewchicken55/NeuralSim), path:line.
 *
 * Parity status of each part (what pins it) is listed in the file header of
 * tests/test_oracle.py; every function below is exercised there.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ===================================================================================== */
/* Primitives                                                                             */
/* ===================================================================================== */

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
 * SC'11).  Round: (L0,R0,L1,R1) -> (hi(M1*R1)^L0^k0? ...) written from the paper's
 * definition: two 32x32->64 products per round, Weyl key schedule. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n1 = lo1;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    uint32_t n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* exp: Cody-Waite reduction x = k ln2 + r (ln2 split in a 15-bit head and a tail so k*head
 * is exact), Taylor polynomial of degree 7 on |r| <= ln2/2 (truncation < 6e-9 relative),
 * scale by 2^k.  Results below 2^-124 flush to 0 (x < -86). */
float orc_expf(float x) {
  if (x < -86.0f) return 0.0f;
  if (x > 88.0f) x = 88.0f;
  const float log2e = 1.44269504088896341f;
  const float ln2_hi = 0.693145751953125f;        /* 0x3F317200 */
  const float ln2_lo = 1.42860682030941723212e-6f;
  float k = rintf(x * log2e);
  float r = x - k * ln2_hi;
  r = r - k * ln2_lo;
  float p = 1.0f / 5040.0f;
  p = p * r + 1.0f / 720.0f;
  p = p * r + 1.0f / 120.0f;
  p = p * r + 1.0f / 24.0f;
  p = p * r + 1.0f / 6.0f;
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  return ldexpf(p, (int)k);
}

/* log for positive normal x: x = 2^e m, m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s),
 * s = (m-1)/(m+1), odd series to s^11; log x = e ln2_hi + (log m + e ln2_lo). */
float orc_logf(float x) {
  uint32_t b = f2u(x);
  int e = (int)((b >> 23) & 0xffu) - 127;
  float m = u2f((b & 0x7fffffu) | 0x3f800000u);
  if (m > 1.41421356f) { m = m * 0.5f; e += 1; }
  float s = (m - 1.0f) / (m + 1.0f);
  float z = s * s;
  float p = 1.0f / 11.0f;
  p = p * z + 1.0f / 9.0f;
  p = p * z + 1.0f / 7.0f;
  p = p * z + 1.0f / 5.0f;
  p = p * z + 1.0f / 3.0f;
  p = p * z;
  float s2 = s + s;
  float lm = s2 * p;
  lm = s2 + lm;
  const float ln2_hi = 0.693145751953125f;
  const float ln2_lo = 1.42860682030941723212e-6f;
  float fe = (float)e;
  float lo = fe * ln2_lo;
  lo = lm + lo;
  float hi = fe * ln2_hi;
  return hi + lo;
}

/* cos(pi x) for x in [0, 2]: x = q/2 + f, |f| <= 1/4; Taylor sin/cos of pi f. */
float orc_cospi(float x) {
  float q = rintf(x + x);
  float f = x - 0.5f * q;
  float y = f * 3.14159265358979f;
  float y2 = y * y;
  float s = 1.0f / 362880.0f;
  s = s * y2 - 1.0f / 5040.0f;
  s = s * y2 + 1.0f / 120.0f;
  s = s * y2 - 1.0f / 6.0f;
  s = s * y2;
  s = s * y;
  s = y + s;
  float c = -1.0f / 3628800.0f;
  c = c * y2 + 1.0f / 40320.0f;
  c = c * y2 - 1.0f / 720.0f;
  c = c * y2 + 1.0f / 24.0f;
  c = c * y2 - 0.5f;
  c = c * y2 + 1.0f;
  int qi = ((int)q) & 3;
  if (qi == 0) return c;
  if (qi == 1) return -s;
  if (qi == 2) return -c;
  return s;
}

/* Box-Muller (Box & Muller 1958): u1 in (0,1], u2 in [0,1); z = sqrt(-2 ln u1) cos(2 pi u2) */
float orc_normal(uint32_t r0, uint32_t r1) {
  float u1 = (float)((r0 >> 8) + 1u) * 5.9604644775390625e-8f; /* 2^-24 */
  float u2 = (float)(r1 >> 8) * 5.9604644775390625e-8f;
  float l = orc_logf(u1);
  l = -2.0f * l;
  float rad = sqrtf(l);
  float c = orc_cospi(u2 + u2);
  return rad * c;
}

double orc_normal64(uint32_t r0, uint32_t r1) {
  double u1 = (double)((r0 >> 8) + 1u) * 5.9604644775390625e-8;
  double u2 = (double)(r1 >> 8) * 5.9604644775390625e-8;
  return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

/* Poisson(lambda) by inversion against integer thresholds thr[k] = floor(CDF(k) 2^32). */
int orc_poisson_table(double lambda, uint32_t* thr, int cap) {
  double p = exp(-lambda), cdf = p;
  int k = 0;
  while (k < cap) {
    double t = floor(cdf * 4294967296.0);
    thr[k] = t >= 4294967295.0 ? 0xffffffffu : (uint32_t)t;
    ++k;
    if (thr[k - 1] == 0xffffffffu) break;
    p = p * lambda / (double)k;
    cdf += p;
  }
  return k;
}

static inline uint32_t poisson_draw(uint32_t x, const uint32_t* thr, int n) {
  uint32_t c = 0;
  while ((int)c < n && x >= thr[c]) ++c;
  return c;
}

#define STREAM_CONN 0x434F4E4Eu
#define STREAM_DRIVE 0x44525645u
#define STREAM_PARAM 0x50415241u
#define NEVER_STEP (-1073741824)

static inline void draw4(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                         uint32_t out[4]) {
  uint32_t ctr[4] = {c0, c1, c2, c3};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  orc_philox4x32_10(ctr, key, out);
}

/* Drive words of neuron i at step t: one Philox draw per four consecutive neurons (group i/4).
   Poisson takes word i%4; Gaussian the pair (i%4)/2, the angle word advanced by three quarter
   turns for odd i (Box-Muller's sine of the pair; the device's drive_stage / drive_draw). */
static inline uint32_t drive_word(uint64_t seed, uint64_t i, uint32_t t) {
  uint32_t rr[4];
  const uint64_t g = i >> 2;
  draw4(seed, (uint32_t)g, t, STREAM_DRIVE, (uint32_t)(g >> 32), rr);
  return rr[i & 3];
}
static inline void drive_pair(uint64_t seed, uint64_t i, uint32_t t, uint32_t* radius,
                              uint32_t* angle) {
  uint32_t rr[4];
  const uint64_t g = i >> 2;
  draw4(seed, (uint32_t)g, t, STREAM_DRIVE, (uint32_t)(g >> 32), rr);
  const uint32_t b = (uint32_t)(i & 2);
  *radius = rr[b];
  *angle = rr[b + 1] + ((i & 1) ? 0xC0000000u : 0u);
}

/* The drive normal of neuron i at step t (FP32, as the mirror and the device draw it). */
float orc_drive_normal(uint64_t seed, uint64_t i, uint32_t t) {
  uint32_t ra, an;
  drive_pair(seed, i, t, &ra, &an);
  return orc_normal(ra, an);
}

/* ===================================================================================== */
/* FP64 restatement of the reference                                                      */
/* ===================================================================================== */

/* LifNeuron::step, src/neuron/lif.rs:15-61 */
int ref_lif_step(double* v, int32_t* refr, double resting, double threshold, double reset,
                 double tau_m, double refractory_period, double input_resistance, double dt,
                 double input_current) {
  if (*refr > 0) { *refr -= 1; return 0; }
  double dv = (-(*v - resting) + input_resistance * input_current) * dt / tau_m;
  *v += dv;
  if (*v >= threshold) {
    *v = reset;
    *refr = (int32_t)round(refractory_period / dt); /* Rust f64::round: half away from 0 */
    return 1;
  }
  return 0;
}

/* IzhikevichNeuron::step, src/neuron/izhikevich.rs:20-52 (spike check first, du on old v) */
int ref_izh_step(double* v, double* u, double a, double b, double c, double d, double dt,
                 double input_current) {
  if (*v >= 30.0) { *v = c; *u += d; return 1; }
  double vv = *v, uu = *u;
  double dv = (0.04 * vv * vv + 5.0 * vv + 140.0 - uu + input_current) * dt;
  double du = a * (b * vv - uu) * dt;
  *v += dv;
  *u += du;
  return 0;
}

/* izhikevich.rs:20-52 with the V integration split in `sub` equal sub-steps of dt/sub (the
 * benchmark network uses two 0.5 ms V sub-steps per 1 ms step); u advances once with the
 * step-start v as the reference's du does.  sub = 1 is ref_izh_step exactly. */
int ref_izh_step_sub(double* v, double* u, double a, double b, double c, double d, double dt,
                     int sub, double input_current) {
  if (*v >= 30.0) { *v = c; *u += d; return 1; }
  double v0 = *v, uu = *u, vv = *v, h = dt / sub;
  for (int s = 0; s < sub; ++s) vv += (0.04 * vv * vv + 5.0 * vv + 140.0 - uu + input_current) * h;
  *u += a * (b * v0 - uu) * dt;
  *v = vv;
  return 0;
}

/* HodgkinHuxleyNeuron rate functions, src/neuron/hodgkin_huxley.rs:20-42.  alpha_m and alpha_n
 * are 0/0 at v = -40 / -55 in the reference; the removable singularity is filled with the
 * limit and its first-order term (|x| < 1e-7). */
double ref_hh_alpha_m(double v) {
  double x = v + 40.0;
  if (fabs(x) < 1e-7) return 1.0 + x / 20.0;
  return 0.1 * (v + 40.0) / (1.0 - exp(-(v + 40.0) / 10.0));
}
double ref_hh_beta_m(double v) { return 4.0 * exp(-(v + 65.0) / 18.0); }
double ref_hh_alpha_h(double v) { return 0.07 * exp(-(v + 65.0) / 20.0); }
double ref_hh_beta_h(double v) { return 1.0 / (1.0 + exp(-(v + 35.0) / 10.0)); }
double ref_hh_alpha_n(double v) {
  double x = v + 55.0;
  if (fabs(x) < 1e-7) return 0.1 + x / 200.0;
  return 0.01 * (v + 55.0) / (1.0 - exp(-(v + 55.0) / 10.0));
}
double ref_hh_beta_n(double v) { return 0.125 * exp(-(v + 65.0) / 80.0); }

/* Gate and membrane update of HodgkinHuxleyNeuron::step, hodgkin_huxley.rs:46-86:
 * gates by Euler on the old V, currents with the new gates and the old V.
 * documented = 0: the code as written, dv = (I - i_na - i_k - i_l)/c_m dt with
 *   i_x = g_x (E_x - V) (hodgkin_huxley.rs:78-84), which pushes V away from every reversal
 *   potential;
 * documented = 1: the membrane equation of the file's doc comment (hodgkin_huxley.rs:6-9),
 *   C dV/dt = g_Na m^3 h (E_Na - V) + g_K n^4 (E_K - V) + g_L (E_L - V) + I. */
double ref_hh_integrate(double* v, double* m, double* h, double* n, double g_na, double g_k,
                        double g_l, double e_na, double e_k, double e_l, double c_m, double dt,
                        double input_current, int documented) {
  double vv = *v;
  double am = ref_hh_alpha_m(vv), bm = ref_hh_beta_m(vv);
  double ah = ref_hh_alpha_h(vv), bh = ref_hh_beta_h(vv);
  double an = ref_hh_alpha_n(vv), bn = ref_hh_beta_n(vv);
  *m += (am * (1.0 - *m) - bm * *m) * dt;
  *h += (ah * (1.0 - *h) - bh * *h) * dt;
  *n += (an * (1.0 - *n) - bn * *n) * dt;
  double i_na = g_na * (*m) * (*m) * (*m) * (*h) * (e_na - vv);
  double n2 = (*n) * (*n);
  double i_k = g_k * n2 * n2 * (e_k - vv);
  double i_l = g_l * (e_l - vv);
  double dv = documented ? (input_current + i_na + i_k + i_l) / c_m * dt
                         : (input_current - i_na - i_k - i_l) / c_m * dt;
  *v += dv;
  return vv;
}

/* HodgkinHuxleyNeuron::step, hodgkin_huxley.rs:46-96 including its V >= -55 -> -70 reset. */
int ref_hh_step(double* v, double* m, double* h, double* n, double g_na, double g_k,
                double g_l, double e_na, double e_k, double e_l, double c_m, double dt,
                double input_current) {
  ref_hh_integrate(v, m, h, n, g_na, g_k, g_l, e_na, e_k, e_l, c_m, dt, input_current, 0);
  if (*v >= -55.0) { *v = -70.0; return 1; }
  return 0;
}

/* Network::finalize, src/network/mod.rs:79-106: out-degree count, prefix sum, stable
 * placement in insertion order.  perm[slot] = insertion index of the synapse in that slot
 * (the reference permutes adjacency_indices only; the build permutes weight/delay with it). */
void ref_finalize(uint64_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                  uint64_t* row_ptr, uint32_t* col, uint64_t* perm) {
  uint64_t* cursor = (uint64_t*)calloc(n + 1, sizeof(uint64_t));
  for (uint64_t k = 0; k < m; ++k)
    if (src[k] < n) cursor[src[k]] += 1;
  uint64_t sum = 0;
  row_ptr[0] = 0;
  for (uint64_t i = 0; i < n; ++i) {
    sum += cursor[i];
    row_ptr[i + 1] = sum;
  }
  for (uint64_t i = 0; i < n; ++i) cursor[i] = row_ptr[i];
  for (uint64_t k = 0; k < m; ++k) {
    if (src[k] < n) {
      uint64_t pos = cursor[src[k]]++;
      col[pos] = dst[k];
      if (perm) perm[pos] = k;
    }
  }
  free(cursor);
}

/* SynapseDynamics::nmda_mg_block, src/synapse/types.rs:88-94 ([Mg] = 1 mM) */
double ref_mg_block(double mg_block_strength, double v) {
  if (mg_block_strength <= 0.0) return 1.0;
  double mg = 1.0;
  return 1.0 / (1.0 + mg / 3.57 * exp(-0.062 * v));
}

/* SynapseDynamics::conductance, types.rs:59-85 (normalised dual exponential) */
double ref_conductance(double rise, double decay, double gmax, double t) {
  if (t < 0.0) return 0.0;
  if (decay <= 0.0 || rise <= 0.0) return 0.0;
  if (fabs(decay - rise) < 1e-12) {
    double norm = rise * exp(-1.0);
    return gmax * t * exp(-t / rise) / norm;
  }
  double t_peak = (rise * decay / (decay - rise)) * log(decay / rise);
  double norm = exp(-t_peak / decay) - exp(-t_peak / rise);
  if (norm <= 0.0) return 0.0;
  double g = (exp(-t / decay) - exp(-t / rise)) / norm;
  return gmax * (g > 0.0 ? g : 0.0);
}

/* StdpTrace, src/synapse/plasticity.rs:144-159 */
void ref_trace_on_pre(double* pre, double* last_pre, double time, double tau_plus) {
  double dt = time - *last_pre;
  *pre = *pre * exp(-dt / tau_plus) + 1.0;
  *last_pre = time;
}
void ref_trace_on_post(double* post, double* last_post, double time, double tau_minus) {
  double dt = time - *last_post;
  *post = *post * exp(-dt / tau_minus) + 1.0;
  *last_post = time;
}
void ref_trace_decay(double* pre, double* post, double dt, double tau_plus, double tau_minus) {
  *pre *= exp(-dt / tau_plus);
  *post *= exp(-dt / tau_minus);
}

/* StdpRule::weight_change, plasticity.rs:31-37 */
double ref_stdp_weight_change(double dt, double a_plus, double a_minus, double tau_plus,
                              double tau_minus) {
  if (dt > 0.0) return a_plus * exp(-dt / tau_plus);
  return -a_minus * exp(dt / tau_minus);
}

/* TripletStdpRule::weight_change, plasticity.rs:75-80 */
double ref_triplet_weight_change(double dt_pre_post, double dt_post_pre, double tau_plus,
                                 double tau_minus, double tau_y, double a_plus,
                                 double a_minus, double a_triplet) {
  double ltp = a_plus * exp(-fabs(dt_pre_post) / tau_plus) +
               a_triplet * exp(-fabs(dt_post_pre) / tau_y);
  double ltd = a_minus * exp(-fabs(dt_pre_post) / tau_minus);
  return dt_pre_post > 0.0 ? ltp : -ltd;
}

/* RateHomeostasis::evaluate, src/evolution/fitness.rs:59-75 (after simulate_ms) */
double ref_rate_homeostasis(double mean_rate, double target_rate, int penalize_silence) {
  double fitness = 1.0 / (1.0 + fabs(mean_rate - target_rate));
  if (penalize_silence && mean_rate < 1.0) fitness *= 0.1;
  if (mean_rate > 500.0) fitness *= 0.1;
  return fitness;
}

static int cmp_u32(const void* a, const void* b) {
  uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
  return x < y ? -1 : x > y;
}

/* engine.rs:496-537 over a given raster, one StdpTrace per plastic synapse.
 * Time: net.time starts at 0 and is advanced by dt before plasticity runs (engine.rs:389),
 * so the trace code sees current_time = net.time after the increment (engine.rs:498). */
void ref_stdp_literal(uint64_t m, const uint32_t* src, const uint32_t* tgt, double* w,
                      const uint8_t* plastic, uint32_t n_steps, const uint64_t* step_ptr,
                      const uint32_t* spikes, uint64_t n_neurons, double dt, double a_plus,
                      double a_minus, double tau_plus, double tau_minus, double w_min,
                      double w_max) {
  double* pre = (double*)calloc(m, sizeof(double));
  double* post = (double*)calloc(m, sizeof(double));
  double* lpre = (double*)malloc(m * sizeof(double));
  double* lpost = (double*)malloc(m * sizeof(double));
  uint8_t* fired = (uint8_t*)calloc(n_neurons, 1);
  for (uint64_t k = 0; k < m; ++k) { lpre[k] = -1e9; lpost[k] = -1e9; }
  double time = 0.0;
  for (uint32_t t = 0; t < n_steps; ++t) {
    time += dt;
    for (uint64_t s = step_ptr[t]; s < step_ptr[t + 1]; ++s) fired[spikes[s]] = 1;
    for (uint64_t k = 0; k < m; ++k) {
      if (!plastic[k]) continue;
      ref_trace_decay(&pre[k], &post[k], dt, tau_plus, tau_minus);
      if (fired[src[k]]) {
        double delta = a_plus * post[k];
        double x = w[k] + delta;
        w[k] = x < w_min ? w_min : (x > w_max ? w_max : x);
        ref_trace_on_pre(&pre[k], &lpre[k], time, tau_plus);
      }
      if (fired[tgt[k]]) {
        double delta = -a_minus * pre[k];
        double x = w[k] + delta;
        w[k] = x < w_min ? w_min : (x > w_max ? w_max : x);
        ref_trace_on_post(&post[k], &lpost[k], time, tau_minus);
      }
    }
    for (uint64_t s = step_ptr[t]; s < step_ptr[t + 1]; ++s) fired[spikes[s]] = 0;
  }
  free(pre); free(post); free(lpre); free(lpost); free(fired);
}

/* Ornstein-Uhlenbeck input current over one step of dt, exact in distribution
 * (dX = (mean - X)/tau dt + sigma sqrt(2/tau) dW): with e = exp(-dt/tau),
 * X' = X e + mean (1 - e) + sigma sqrt(1 - e^2) z, z ~ N(0, 1).  The stationary law is
 * N(mean, sigma^2) with lag-k autocorrelation e^k.  (The 200k-neuron HH workload: mean 6,
 * sigma 3, tau 5 ms; the reference has no OU source -- engine.rs:273-287 injects uniform noise.) */
double ref_ou_update(double x, double mean, double sigma, double tau, double dt, double z) {
  double e = exp(-dt / tau);
  return x * e + mean * (1.0 - e) + sigma * sqrt(1.0 - e * e) * z;
}

/* ---- FP64 network run ---------------------------------------------------------------- */
struct ref_net {
  nsx_config cfg;
  uint64_t n, m, t;
  int ch;
  uint64_t* row_ptr;
  uint32_t* col;
  uint8_t* delay;
  double* w;
  double* ring; /* D*ch*n */
  double *v, *u, *gm, *gh, *gn, *s0, *s1, *s2, *drive, *c, *d;
  int32_t* refr;
  uint32_t* pthr;
  int npthr;
  /* literal per-synapse STDP traces */
  double *pre, *post, *lpre, *lpost;
  uint8_t* fired;
  uint32_t* src_of;
  double time;
};

static int channels_of(const nsx_config* c) {
  return c->model == NSX_MODEL_IZHIKEVICH ? 1 : 2;
}

static void izh_params(const nsx_config* cfg, uint64_t i, float* c, float* d) {
  uint32_t r[4];
  draw4(cfg->seed, (uint32_t)i, 0u, STREAM_PARAM, (uint32_t)(i >> 32), r);
  float x = (float)(r[0] >> 8) * 5.9604644775390625e-8f;
  float x2 = x * x;
  int ty = i < cfg->n_exc ? 0 : 1;
  float cj = cfg->izh_c_jitter * x2;
  *c = cfg->izh_c[ty] + cj;
  float dj = cfg->izh_d_jitter_frac * x2;
  dj = 1.0f - dj;
  *d = cfg->izh_d[ty] * dj;
}

ref_net* ref_net_create(const nsx_config* cfg, const uint64_t* row_ptr, const uint32_t* col,
                        const float* w, const uint8_t* delay) {
  ref_net* r = (ref_net*)calloc(1, sizeof(ref_net));
  r->cfg = *cfg;
  r->n = cfg->n_neurons;
  r->m = row_ptr[r->n];
  r->ch = channels_of(cfg);
  uint64_t n = r->n, m = r->m;
  r->row_ptr = (uint64_t*)malloc((n + 1) * 8);
  memcpy(r->row_ptr, row_ptr, (n + 1) * 8);
  r->col = (uint32_t*)malloc(m * 4 + 4);
  memcpy(r->col, col, m * 4);
  r->delay = (uint8_t*)malloc(m + 1);
  memcpy(r->delay, delay, m);
  r->w = (double*)malloc(m * 8 + 8);
  for (uint64_t k = 0; k < m; ++k) r->w[k] = w[k];
  r->src_of = (uint32_t*)malloc(m * 4 + 4);
  for (uint64_t i = 0; i < n; ++i)
    for (uint64_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) r->src_of[k] = (uint32_t)i;
  r->ring = (double*)calloc((size_t)cfg->ring_slots * r->ch * n, 8);
  double** arrs[] = {&r->v, &r->u, &r->gm, &r->gh, &r->gn, &r->s0, &r->s1, &r->s2, &r->drive,
                     &r->c, &r->d};
  for (size_t a = 0; a < sizeof(arrs) / sizeof(arrs[0]); ++a) *arrs[a] = (double*)calloc(n, 8);
  r->refr = (int32_t*)calloc(n, 4);
  r->pthr = (uint32_t*)calloc(64, 4);
  if (cfg->drive == NSX_DRIVE_POISSON)
    r->npthr = orc_poisson_table((double)cfg->drive_inputs * cfg->drive_rate_hz * cfg->dt_ms / 1000.0,
                                 r->pthr, 64);
  for (uint64_t i = 0; i < n; ++i) {
    if (cfg->model == NSX_MODEL_LIF) {
      r->v[i] = cfg->lif_v_rest;
    } else if (cfg->model == NSX_MODEL_IZHIKEVICH) {
      float c, d;
      izh_params(cfg, i, &c, &d);
      r->c[i] = c;
      r->d[i] = d;
      r->v[i] = c; /* NeuronState::new: membrane = resting_potential() = c (mod.rs:45-60,97-103) */
      r->u[i] = 0.0;
    } else {
      r->v[i] = -65.0; /* HodgkinHuxleyNeuron::reset_state, hodgkin_huxley.rs:98-114 */
      r->gm[i] = 0.05; r->gh[i] = 0.6; r->gn[i] = 0.32;
      r->drive[i] = cfg->drive == NSX_DRIVE_OU ? cfg->drive_mean : 0.0;
    }
  }
  if (cfg->plasticity == NSX_PLAST_PAIR) {
    r->pre = (double*)calloc(m + 1, 8);
    r->post = (double*)calloc(m + 1, 8);
    r->lpre = (double*)malloc(m * 8 + 8);
    r->lpost = (double*)malloc(m * 8 + 8);
    for (uint64_t k = 0; k < m; ++k) { r->lpre[k] = -1e9; r->lpost[k] = -1e9; }
    r->fired = (uint8_t*)calloc(n, 1);
  }
  return r;
}

void ref_net_destroy(ref_net* r) {
  if (!r) return;
  free(r->row_ptr); free(r->col); free(r->delay); free(r->w); free(r->src_of); free(r->ring);
  free(r->v); free(r->u); free(r->gm); free(r->gh); free(r->gn); free(r->s0); free(r->s1);
  free(r->s2); free(r->drive); free(r->c); free(r->d); free(r->refr); free(r->pthr);
  free(r->pre); free(r->post); free(r->lpre); free(r->lpost); free(r->fired);
  free(r);
}

uint64_t ref_net_run(ref_net* r, uint32_t n_steps, uint32_t* sp_step, uint32_t* sp_id,
                     uint64_t cap) {
  const nsx_config* cfg = &r->cfg;
  uint64_t n = r->n, count = 0;
  double dt = cfg->dt_ms;
  int D = (int)cfg->ring_slots, ch = r->ch;
  uint32_t* spk = (uint32_t*)malloc(n * 4 + 4);
  for (uint32_t step = 0; step < n_steps; ++step, ++r->t) {
    uint64_t t = r->t;
    int slot = (int)(t % (uint64_t)D);
    uint64_t ns = 0;
    for (uint64_t i = 0; i < n; ++i) {
      double a0 = r->ring[((size_t)slot * ch + 0) * n + i];
      double a1 = ch > 1 ? r->ring[((size_t)slot * ch + 1) * n + i] : 0.0;
      r->ring[((size_t)slot * ch + 0) * n + i] = 0.0;
      if (ch > 1) r->ring[((size_t)slot * ch + 1) * n + i] = 0.0;
      double z = 0.0, pois = 0.0;
      if (cfg->drive == NSX_DRIVE_POISSON) {
        pois = (double)poisson_draw(drive_word(cfg->seed, i, (uint32_t)t), r->pthr, r->npthr) *
               cfg->drive_weight;
      } else if (cfg->drive != NSX_DRIVE_NONE) {
        uint32_t ra, an;
        drive_pair(cfg->seed, i, (uint32_t)t, &ra, &an);
        z = orc_normal64(ra, an);
      }
      int spiked = 0;
      if (cfg->model == NSX_MODEL_LIF) {
        a0 += pois;
        r->s0[i] = r->s0[i] * exp(-dt / cfg->tau_exc_ms) + a0;
        r->s1[i] = r->s1[i] * exp(-dt / cfg->tau_inh_ms) + a1;
        spiked = ref_lif_step(&r->v[i], &r->refr[i], cfg->lif_v_rest, cfg->lif_v_th,
                              cfg->lif_v_reset, cfg->lif_tau_m, cfg->lif_t_ref_ms, cfg->lif_r,
                              dt, r->s0[i] + r->s1[i]);
      } else if (cfg->model == NSX_MODEL_IZHIKEVICH) {
        double I = a0 + pois;
        if (cfg->drive == NSX_DRIVE_GAUSSIAN) I += cfg->drive_mean + cfg->drive_sigma * z;
        int ty = i < cfg->n_exc ? 0 : 1;
        spiked = ref_izh_step_sub(&r->v[i], &r->u[i], cfg->izh_a[ty], cfg->izh_b[ty], r->c[i],
                                  r->d[i], dt, (int)cfg->v_substeps, I);
      } else {
        double v = r->v[i];
        a0 += pois;
        r->s0[i] = r->s0[i] * exp(-dt / cfg->tau_exc_ms) + a0 * cfg->gmax_ampa;
        r->s2[i] = r->s2[i] * exp(-dt / cfg->tau_nmda_ms);
        if (cfg->mg_mM > 0.0f)
          r->s2[i] += a0 * cfg->gmax_nmda / (1.0 + cfg->mg_mM / 3.57 * exp(-0.062 * v));
        r->s1[i] = r->s1[i] * exp(-dt / cfg->tau_inh_ms) + fabs(a1) * cfg->gmax_gaba;
        double isyn = r->s0[i] * (cfg->e_exc - v) + r->s2[i] * (cfg->e_exc - v) +
                      r->s1[i] * (cfg->e_inh - v);
        double iext = 0.0;
        if (cfg->drive == NSX_DRIVE_OU) {
          iext = r->drive[i];
          r->drive[i] = ref_ou_update(r->drive[i], cfg->drive_mean, cfg->drive_sigma,
                                      cfg->drive_tau_ms, dt, z);
        } else if (cfg->drive == NSX_DRIVE_GAUSSIAN) {
          iext = cfg->drive_mean + cfg->drive_sigma * z;
        }
        const uint32_t sub = cfg->hh_substeps ? cfg->hh_substeps : 1u;
        for (uint32_t k = 0; k < sub; ++k) /* input current held over the sub-steps */
          ref_hh_integrate(&r->v[i], &r->gm[i], &r->gh[i], &r->gn[i], cfg->hh_g_na, cfg->hh_g_k,
                           cfg->hh_g_l, cfg->hh_e_na, cfg->hh_e_k, cfg->hh_e_l, cfg->hh_c_m,
                           dt / sub, iext + isyn,
                           cfg->hh_spike_mode != NSX_HH_SPIKE_REFERENCE_RESET);
        if (cfg->hh_spike_mode == NSX_HH_SPIKE_REFERENCE_RESET) {
          if (r->v[i] >= -55.0) { r->v[i] = -70.0; spiked = 1; }
        } else {
          spiked = v < cfg->hh_spike_v && r->v[i] >= cfg->hh_spike_v;
        }
      }
      if (spiked) {
        spk[ns++] = (uint32_t)i;
        if (count < cap) { sp_step[count] = (uint32_t)t; sp_id[count] = (uint32_t)i; }
        count++;
      }
    }
    /* propagation with delays: arrival at step t + d */
    for (uint64_t s = 0; s < ns; ++s) {
      uint32_t i = spk[s];
      int chn = (ch > 1 && i >= cfg->n_exc) ? 1 : 0;
      for (uint64_t k = r->row_ptr[i]; k < r->row_ptr[i + 1]; ++k) {
        int sl = (int)((t + r->delay[k]) % (uint64_t)D);
        r->ring[((size_t)sl * ch + chn) * n + r->col[k]] += r->w[k];
      }
    }
    /* literal per-synapse STDP (engine.rs:496-537), E->E synapses plastic */
    if (cfg->plasticity == NSX_PLAST_PAIR) {
      r->time += dt;
      for (uint64_t s = 0; s < ns; ++s) r->fired[spk[s]] = 1;
      for (uint64_t k = 0; k < r->m; ++k) {
        if (r->src_of[k] >= cfg->n_exc || r->col[k] >= cfg->n_exc) continue;
        ref_trace_decay(&r->pre[k], &r->post[k], dt, cfg->tau_plus_ms, cfg->tau_minus_ms);
        if (r->fired[r->src_of[k]]) {
          double x = r->w[k] + cfg->a_plus * r->post[k];
          r->w[k] = x < cfg->w_min ? cfg->w_min : (x > cfg->w_max ? cfg->w_max : x);
          ref_trace_on_pre(&r->pre[k], &r->lpre[k], r->time, cfg->tau_plus_ms);
        }
        if (r->fired[r->col[k]]) {
          double x = r->w[k] + (-cfg->a_minus * r->pre[k]);
          r->w[k] = x < cfg->w_min ? cfg->w_min : (x > cfg->w_max ? cfg->w_max : x);
          ref_trace_on_post(&r->post[k], &r->lpost[k], r->time, cfg->tau_minus_ms);
        }
      }
      for (uint64_t s = 0; s < ns; ++s) r->fired[spk[s]] = 0;
    }
  }
  free(spk);
  return count;
}

void ref_net_state(ref_net* r, double* v, double* u, double* m, double* h, double* n) {
  size_t b = r->n * 8;
  if (v) memcpy(v, r->v, b);
  if (u) memcpy(u, r->u, b);
  if (m) memcpy(m, r->gm, b);
  if (h) memcpy(h, r->gh, b);
  if (n) memcpy(n, r->gn, b);
}

void ref_net_weights(ref_net* r, double* w) { memcpy(w, r->w, r->m * 8); }

/* The OU drive current of every neuron (HH networks with an OU drive). */
void ref_net_drive(ref_net* r, double* x) { memcpy(x, r->drive, r->n * 8); }

/* ===================================================================================== */
/* FP32 mirror of the device engine                                                       */
/* ===================================================================================== */
struct mir_net {
  nsx_config cfg;
  uint64_t n, m, t;
  int ch, D, S;
  uint64_t* row_ptr;
  uint32_t* col;
  float* w;
  uint8_t* delay;
  /* E->E transpose for post-spike depression: slots and sources per target */
  uint64_t* csc_ptr;
  uint64_t* csc_slot;
  uint32_t* csc_src;
  int64_t* ring;
  float *v, *u, *gm, *gh, *gn, *s0, *s1, *s2, *drive, *c, *d;
  int32_t *refr, *last;
  uint32_t* count;
  float *pre_v, *post_v;
  int32_t *pre_s, *post_s;
  float *ptab, *mtab, *ytab;
  int ptab_n, mtab_n, ytab_n;
  uint32_t pthr[64];
  int npthr;
  /* precomputed float constants (host double -> float, as the device does) */
  float dec_e, dec_i, dec_n, lif_k, ou_k1, ou_k2, ou_k3, izh_h, dtf, hh_kv, mgc, scale_in;
  float hh_dts;
  uint32_t hh_sub;
  int32_t refr_steps;
  double scale_out;
};

static float* trace_table(double dt, double tau, int* len) {
  int n = (int)ceil(104.0 * tau / dt) + 2;
  float* t = (float*)malloc((size_t)n * 4);
  for (int k = 0; k < n; ++k) t[k] = (float)exp(-(double)k * dt / tau);
  *len = n;
  return t;
}

static inline float tab(const float* t, int n, int64_t k) {
  return (k >= 0 && k < n) ? t[k] : 0.0f;
}

mir_net* mir_create(const nsx_config* cfg) {
  mir_net* r = (mir_net*)calloc(1, sizeof(mir_net));
  r->cfg = *cfg;
  r->n = cfg->n_neurons;
  r->ch = channels_of(cfg);
  r->D = (int)cfg->ring_slots;
  r->S = (int)cfg->frac_bits;
  uint64_t n = r->n;
  r->ring = (int64_t*)calloc((size_t)r->D * r->ch * n, 8);
  float** arrs[] = {&r->v, &r->u, &r->gm, &r->gh, &r->gn, &r->s0, &r->s1, &r->s2, &r->drive,
                    &r->c, &r->d, &r->pre_v, &r->post_v};
  for (size_t a = 0; a < sizeof(arrs) / sizeof(arrs[0]); ++a) *arrs[a] = (float*)calloc(n, 4);
  r->refr = (int32_t*)calloc(n, 4);
  r->last = (int32_t*)malloc(n * 4 + 4);
  r->pre_s = (int32_t*)malloc(n * 4 + 4);
  r->post_s = (int32_t*)malloc(n * 4 + 4);
  r->count = (uint32_t*)calloc(n, 4);
  double dt = cfg->dt_ms;
  r->dtf = (float)dt;
  r->dec_e = (float)exp(-dt / cfg->tau_exc_ms);
  r->dec_i = (float)exp(-dt / cfg->tau_inh_ms);
  r->dec_n = (float)exp(-dt / cfg->tau_nmda_ms);
  r->lif_k = (float)(dt / cfg->lif_tau_m);
  r->refr_steps = (int32_t)round((double)cfg->lif_t_ref_ms / dt);
  r->izh_h = (float)(dt / (cfg->v_substeps ? cfg->v_substeps : 1));
  r->hh_sub = cfg->hh_substeps ? cfg->hh_substeps : 1u;
  r->hh_dts = (float)(dt / r->hh_sub);
  r->hh_kv = (float)(dt / r->hh_sub / cfg->hh_c_m);
  r->mgc = (float)(cfg->mg_mM / 3.57);
  r->scale_in = ldexpf(1.0f, r->S);
  r->scale_out = ldexp(1.0, -r->S);
  if (cfg->drive == NSX_DRIVE_OU) {
    double e = exp(-dt / cfg->drive_tau_ms);
    r->ou_k1 = (float)e;
    r->ou_k2 = (float)(cfg->drive_mean * (1.0 - e));
    r->ou_k3 = (float)(cfg->drive_sigma * sqrt(1.0 - e * e));
  }
  if (cfg->drive == NSX_DRIVE_POISSON)
    r->npthr = orc_poisson_table((double)cfg->drive_inputs * cfg->drive_rate_hz * dt / 1000.0,
                                 r->pthr, 64);
  if (cfg->plasticity != NSX_PLAST_NONE) {
    r->ptab = trace_table(dt, cfg->tau_plus_ms, &r->ptab_n);
    r->mtab = trace_table(dt, cfg->tau_minus_ms, &r->mtab_n);
  }
  if (cfg->plasticity == NSX_PLAST_TRIPLET) r->ytab = trace_table(dt, cfg->tau_y_ms, &r->ytab_n);
  nsx_state_view none;
  memset(&none, 0, sizeof(none));
  for (uint64_t i = 0; i < n; ++i) {
    r->last[i] = NEVER_STEP;
    r->pre_s[i] = NEVER_STEP;
    r->post_s[i] = NEVER_STEP;
    if (cfg->model == NSX_MODEL_LIF) {
      r->v[i] = cfg->lif_v_rest;
    } else if (cfg->model == NSX_MODEL_IZHIKEVICH) {
      izh_params(cfg, i, &r->c[i], &r->d[i]);
      r->v[i] = r->c[i];
      r->u[i] = 0.0f;
    } else {
      r->v[i] = -65.0f;
      r->gm[i] = 0.05f; r->gh[i] = 0.6f; r->gn[i] = 0.32f;
      r->drive[i] = cfg->drive == NSX_DRIVE_OU ? cfg->drive_mean : 0.0f;
    }
  }
  return r;
}

void mir_destroy(mir_net* r) {
  if (!r) return;
  free(r->row_ptr); free(r->col); free(r->w); free(r->delay); free(r->csc_ptr);
  free(r->csc_slot); free(r->csc_src); free(r->ring);
  free(r->v); free(r->u); free(r->gm); free(r->gh); free(r->gn); free(r->s0); free(r->s1);
  free(r->s2); free(r->drive); free(r->c); free(r->d); free(r->pre_v); free(r->post_v);
  free(r->refr); free(r->last); free(r->pre_s); free(r->post_s); free(r->count);
  free(r->ptab); free(r->mtab); free(r->ytab);
  free(r);
}

static void build_csc(mir_net* r) {
  uint64_t ne = r->cfg.n_exc;
  free(r->csc_ptr); free(r->csc_slot); free(r->csc_src);
  r->csc_ptr = (uint64_t*)calloc(ne + 1, 8);
  uint64_t tot = 0;
  for (uint64_t i = 0; i < ne; ++i)
    for (uint64_t k = r->row_ptr[i]; k < r->row_ptr[i + 1]; ++k)
      if (r->col[k] < ne) r->csc_ptr[r->col[k] + 1]++;
  for (uint64_t j = 0; j < ne; ++j) r->csc_ptr[j + 1] += r->csc_ptr[j];
  tot = r->csc_ptr[ne];
  r->csc_slot = (uint64_t*)malloc(tot * 8 + 8);
  r->csc_src = (uint32_t*)malloc(tot * 4 + 4);
  uint64_t* cur = (uint64_t*)malloc(ne * 8 + 8);
  memcpy(cur, r->csc_ptr, ne * 8);
  for (uint64_t i = 0; i < ne; ++i)
    for (uint64_t k = r->row_ptr[i]; k < r->row_ptr[i + 1]; ++k)
      if (r->col[k] < ne) {
        uint64_t p = cur[r->col[k]]++;
        r->csc_slot[p] = k;
        r->csc_src[p] = (uint32_t)i;
      }
  free(cur);
}

int mir_build_csr(mir_net* r, const uint32_t* src, const uint32_t* dst, const float* w,
                  const uint8_t* delay, uint64_t m) {
  uint64_t n = r->n;
  for (uint64_t k = 0; k < m; ++k)
    if (src[k] >= n || dst[k] >= n || delay[k] < 1 || delay[k] >= (uint32_t)r->D) return -4;
  free(r->row_ptr); free(r->col); free(r->w); free(r->delay);
  r->row_ptr = (uint64_t*)malloc((n + 1) * 8);
  r->col = (uint32_t*)malloc(m * 4 + 4);
  r->w = (float*)malloc(m * 4 + 4);
  r->delay = (uint8_t*)malloc(m + 1);
  uint64_t* perm = (uint64_t*)malloc(m * 8 + 8);
  ref_finalize(n, m, src, dst, r->row_ptr, r->col, perm);
  for (uint64_t s = 0; s < m; ++s) {
    r->w[s] = w[perm[s]];
    r->delay[s] = delay[perm[s]];
  }
  free(perm);
  r->m = m;
  if (r->cfg.plasticity != NSX_PLAST_NONE) build_csc(r);
  return 0;
}

void mir_generate_list(const nsx_config* cfg, uint32_t k_in, float w_exc_lo, float w_exc_hi,
                       float w_inh_lo, float w_inh_hi, uint32_t d_lo, uint32_t d_hi,
                       uint32_t* src, uint32_t* dst, float* w, uint8_t* delay) {
  uint64_t n = cfg->n_neurons;
  for (uint64_t j = 0; j < n; ++j) {
    for (uint32_t k = 0; k < k_in; ++k) {
      uint64_t idx = j * k_in + k;
      uint32_t rr[4];
      draw4(cfg->seed, (uint32_t)j, k, STREAM_CONN, (uint32_t)(j >> 32), rr);
      uint32_t s = (uint32_t)(((uint64_t)rr[0] * (n - 1)) >> 32);
      if (s >= j) s += 1;
      float uu = (float)(rr[1] >> 8) * 5.9604644775390625e-8f;
      int exc = s < cfg->n_exc;
      float lo = exc ? w_exc_lo : w_inh_lo, hi = exc ? w_exc_hi : w_inh_hi;
      float span = hi - lo;
      float x = span * uu;
      src[idx] = s;
      dst[idx] = (uint32_t)j;
      w[idx] = lo + x;
      delay[idx] = (uint8_t)(d_lo + (uint32_t)(((uint64_t)rr[2] * (d_hi - d_lo + 1)) >> 32));
    }
  }
}

int mir_generate_fixed_indegree(mir_net* r, uint32_t k_in, float w_exc_lo, float w_exc_hi,
                                float w_inh_lo, float w_inh_hi, uint32_t d_lo, uint32_t d_hi) {
  uint64_t m = r->n * k_in;
  uint32_t* src = (uint32_t*)malloc(m * 4 + 4);
  uint32_t* dst = (uint32_t*)malloc(m * 4 + 4);
  float* w = (float*)malloc(m * 4 + 4);
  uint8_t* d = (uint8_t*)malloc(m + 1);
  mir_generate_list(&r->cfg, k_in, w_exc_lo, w_exc_hi, w_inh_lo, w_inh_hi, d_lo, d_hi, src,
                    dst, w, d);
  int rc = mir_build_csr(r, src, dst, w, d, m);
  free(src); free(dst); free(w); free(d);
  return rc;
}

uint64_t mir_num_synapses(mir_net* r) { return r->m; }
uint64_t mir_step_index(mir_net* r) { return r->t; }

void mir_read_csr(mir_net* r, uint64_t* row_ptr, uint32_t* col, float* w, uint8_t* delay) {
  if (row_ptr) memcpy(row_ptr, r->row_ptr, (r->n + 1) * 8);
  if (col) memcpy(col, r->col, r->m * 4);
  if (w) memcpy(w, r->w, r->m * 4);
  if (delay) memcpy(delay, r->delay, r->m);
}

void mir_stimulate(mir_net* r, uint64_t neuron, float current) {
  int slot = (int)(r->t % (uint64_t)r->D);
  r->ring[(size_t)slot * r->ch * r->n + neuron] += (int64_t)llrint((double)current * ldexp(1.0, r->S));
}

/* float HH rate functions in the device's op order (multiplications by reciprocal
 * constants instead of divisions, removable singularities filled for |x| < 1e-3) */
static inline float mir_ratio(float x, float scale_lin, float k) {
  /* scale * x / (1 - exp(-x/10)); near 0: scale*10*(1 + x/20) */
  if (fabsf(x) < 1e-3f) {
    float y = x * 0.05f;
    y = 1.0f + y;
    return k * y;
  }
  float e = orc_expf(x * -0.1f);
  float den = 1.0f - e;
  float num = scale_lin * x;
  return num / den;
}

/* One Euler sub-step with I held: gates at the old V, currents with the new gates and the old
 * V (hodgkin_huxley.rs:46-86); literal = the code's current signs, else the documented
 * equation. */
typedef struct {
  float am, bm, ah, bh, an, bn;
} mir_rates;

static mir_rates mir_hh_rates(float v) {
  mir_rates r;
  r.am = mir_ratio(v + 40.0f, 0.1f, 1.0f);
  r.bm = 4.0f * orc_expf((v + 65.0f) * (-1.0f / 18.0f));
  r.ah = 0.07f * orc_expf((v + 65.0f) * -0.05f);
  float eh = orc_expf((v + 35.0f) * -0.1f);
  eh = 1.0f + eh;
  r.bh = 1.0f / eh;
  r.an = mir_ratio(v + 55.0f, 0.01f, 0.1f);
  r.bn = 0.125f * orc_expf((v + 65.0f) * -0.0125f);
  return r;
}

static void mir_hh_apply(const nsx_config* cfg, const mir_rates* r, float* pv, float* pm,
                         float* ph, float* pn, float I, float dts, float kv, int literal) {
  const float v = *pv;
  float m = *pm, h = *ph, nn = *pn;
  float t1, t2;
  t1 = r->am * (1.0f - m); t2 = r->bm * m; t1 = t1 - t2; t1 = t1 * dts; m = m + t1;
  t1 = r->ah * (1.0f - h); t2 = r->bh * h; t1 = t1 - t2; t1 = t1 * dts; h = h + t1;
  t1 = r->an * (1.0f - nn); t2 = r->bn * nn; t1 = t1 - t2; t1 = t1 * dts; nn = nn + t1;
  float m3 = m * m; m3 = m3 * m;
  float ina = cfg->hh_g_na * m3; ina = ina * h; ina = ina * (cfg->hh_e_na - v);
  float n2 = nn * nn; float n4 = n2 * n2;
  float ik = cfg->hh_g_k * n4; ik = ik * (cfg->hh_e_k - v);
  float il = cfg->hh_g_l * (cfg->hh_e_l - v);
  float acc;
  if (literal) {
    acc = I - ina; acc = acc - ik; acc = acc - il;
  } else {
    acc = I + ina; acc = acc + ik; acc = acc + il;
  }
  acc = acc * kv;
  *pv = v + acc; *pm = m; *ph = h; *pn = nn;
}

static void mir_hh_substep(const nsx_config* cfg, float* pv, float* pm, float* ph, float* pn,
                           float I, float dts, float kv, int literal) {
  const mir_rates r = mir_hh_rates(*pv);
  mir_hh_apply(cfg, &r, pv, pm, ph, pn, I, dts, kv, literal);
}

/* FP32 mirrors of one network step's membrane integration for a single neuron: `substeps` plain
 * Euler sub-steps, or the surrogate's frozen-rate sub-steps (rates at the step-start V). */
void orc_hh_plain(const nsx_config* cfg, float* v, float* m, float* h, float* n, float I,
                  uint32_t substeps) {
  const double dt = cfg->dt_ms;
  const float dts = (float)(dt / substeps), kv = (float)(dt / substeps / cfg->hh_c_m);
  for (uint32_t k = 0; k < substeps; ++k) mir_hh_substep(cfg, v, m, h, n, I, dts, kv, 0);
}

void orc_hh_frozen(const nsx_config* cfg, float* v, float* m, float* h, float* n, float I,
                   uint32_t substeps) {
  const double dt = cfg->dt_ms;
  const float dts = (float)(dt / substeps), kv = (float)(dt / substeps / cfg->hh_c_m);
  const mir_rates r = mir_hh_rates(*v);
  for (uint32_t k = 0; k < substeps; ++k) mir_hh_apply(cfg, &r, v, m, h, n, I, dts, kv, 0);
}

static int mir_hh_update(const mir_net* r, uint64_t i, float a0, float a1, float z) {
  const nsx_config* cfg = &r->cfg;
  float v = r->v[i];
  float ga = r->s0[i] * r->dec_e;
  float xa = a0 * cfg->gmax_ampa;
  ga = ga + xa;
  float gn = r->s2[i] * r->dec_n;
  if (cfg->mg_mM > 0.0f) {
    float e = orc_expf(-0.062f * v);
    float bl = r->mgc * e;
    bl = 1.0f + bl;
    float xn = a0 * cfg->gmax_nmda;
    xn = xn / bl;
    gn = gn + xn;
  }
  float gg = r->s1[i] * r->dec_i;
  float xg = fabsf(a1) * cfg->gmax_gaba;
  gg = gg + xg;
  float de = cfg->e_exc - v, di = cfg->e_inh - v;
  float isyn = ga * de;
  float tn = gn * de;
  isyn = isyn + tn;
  float tg = gg * di;
  isyn = isyn + tg;
  float iext = 0.0f;
  if (cfg->drive == NSX_DRIVE_OU) {
    iext = r->drive[i];
    float o = iext * r->ou_k1;
    o = o + r->ou_k2;
    float q = r->ou_k3 * z;
    r->drive[i] = o + q;
  } else if (cfg->drive == NSX_DRIVE_GAUSSIAN) {
    float q = cfg->drive_sigma * z;
    iext = cfg->drive_mean + q;
  }
  float I = iext + isyn; /* held over the step's Euler sub-steps */
  float m = r->gm[i], h = r->gh[i], nn = r->gn[i];
  float vn = v;
  const int literal = cfg->hh_spike_mode == NSX_HH_SPIKE_REFERENCE_RESET;
  for (uint32_t k = 0; k < r->hh_sub; ++k)
    mir_hh_substep(cfg, &vn, &m, &h, &nn, I, r->hh_dts, r->hh_kv, literal);
  int spiked;
  if (cfg->hh_spike_mode == NSX_HH_SPIKE_REFERENCE_RESET) {
    spiked = vn >= -55.0f;
    if (spiked) vn = -70.0f;
  } else {
    spiked = (v < cfg->hh_spike_v) && (vn >= cfg->hh_spike_v);
  }
  r->v[i] = vn; r->gm[i] = m; r->gh[i] = h; r->gn[i] = nn;
  r->s0[i] = ga; r->s1[i] = gg; r->s2[i] = gn;
  return spiked;
}

uint64_t mir_run(mir_net* r, uint32_t n_steps, uint32_t* sp_step, uint32_t* sp_id,
                 uint64_t cap, uint64_t* events_out) {
  const nsx_config* cfg = &r->cfg;
  uint64_t n = r->n, count = 0, events = 0;
  int D = r->D, ch = r->ch;
  uint64_t ne = cfg->n_exc;
  uint32_t* spk = (uint32_t*)malloc(n * 4 + 4);
  const int pair = cfg->plasticity == NSX_PLAST_PAIR;
  const int trip = cfg->plasticity == NSX_PLAST_TRIPLET;
  for (uint32_t step = 0; step < n_steps; ++step, ++r->t) {
    uint64_t t = r->t;
    int32_t ti = (int32_t)t;
    int slot = (int)(t % (uint64_t)D);
    uint64_t ns = 0;
    /* ---- neuron update ---- */
    for (uint64_t i = 0; i < n; ++i) {
      int64_t* p0 = &r->ring[((size_t)slot * ch + 0) * n + i];
      float a0 = (float)((double)(*p0) * r->scale_out);
      *p0 = 0;
      float a1 = 0.0f;
      if (ch > 1) {
        int64_t* p1 = &r->ring[((size_t)slot * ch + 1) * n + i];
        a1 = (float)((double)(*p1) * r->scale_out);
        *p1 = 0;
      }
      float z = 0.0f;
      if (cfg->drive == NSX_DRIVE_POISSON) {
        float k = (float)poisson_draw(drive_word(cfg->seed, i, (uint32_t)t), r->pthr, r->npthr);
        float x = k * cfg->drive_weight;
        a0 = a0 + x;
      } else if (cfg->drive != NSX_DRIVE_NONE) {
        uint32_t ra, an;
        drive_pair(cfg->seed, i, (uint32_t)t, &ra, &an);
        z = orc_normal(ra, an);
      }
      int spiked = 0;
      if (cfg->model == NSX_MODEL_LIF) {
        float ie = r->s0[i] * r->dec_e;
        ie = ie + a0;
        float ii = r->s1[i] * r->dec_i;
        ii = ii + a1;
        r->s0[i] = ie;
        r->s1[i] = ii;
        if (r->refr[i] > 0) {
          r->refr[i] -= 1;
        } else {
          float I = ie + ii;
          float x = r->v[i] - cfg->lif_v_rest;
          x = -x;
          float y = cfg->lif_r * I;
          x = x + y;
          x = x * r->lif_k;
          float vn = r->v[i] + x;
          if (vn >= cfg->lif_v_th) {
            vn = cfg->lif_v_reset;
            r->refr[i] = r->refr_steps;
            spiked = 1;
          }
          r->v[i] = vn;
        }
      } else if (cfg->model == NSX_MODEL_IZHIKEVICH) {
        float I = a0;
        if (cfg->drive == NSX_DRIVE_GAUSSIAN) {
          float q = cfg->drive_sigma * z;
          q = cfg->drive_mean + q;
          I = I + q;
        }
        float v = r->v[i], u = r->u[i];
        if (v >= 30.0f) {
          v = r->c[i];
          u = u + r->d[i];
          spiked = 1;
        } else {
          int ty = i < ne ? 0 : 1;
          float a = cfg->izh_a[ty], b = cfg->izh_b[ty];
          float v0 = v;
          for (uint32_t s = 0; s < cfg->v_substeps; ++s) {
            float t1 = 0.04f * v;
            t1 = t1 * v;
            float t2 = 5.0f * v;
            float acc = t1 + t2;
            acc = acc + 140.0f;
            acc = acc - u;
            acc = acc + I;
            acc = acc * r->izh_h;
            v = v + acc;
          }
          float t3 = b * v0;
          t3 = t3 - u;
          t3 = a * t3;
          t3 = t3 * r->dtf;
          u = u + t3;
        }
        r->v[i] = v;
        r->u[i] = u;
      } else {
        spiked = mir_hh_update(r, i, a0, a1, z);
      }
      if (spiked) {
        r->count[i] += 1;
        if (trip && i < ne) r->post_s[i] = r->last[i]; /* previous post spike */
        r->last[i] = ti;
        spk[ns++] = (uint32_t)i;
        if (count < cap) { sp_step[count] = (uint32_t)t; sp_id[count] = (uint32_t)i; }
        count++;
        if (pair && i < ne) {
          float p = tab(r->ptab, r->ptab_n, (int64_t)ti - r->pre_s[i]);
          float x = r->pre_v[i] * p;
          x = x * p;
          r->pre_v[i] = x + 1.0f;
          r->pre_s[i] = ti;
        }
      }
    }
    /* ---- propagation (+ potentiation on rows of spiking excitatory sources) ---- */
    for (uint64_t s = 0; s < ns; ++s) {
      uint32_t i = spk[s];
      int chn = (ch > 1 && i >= ne) ? 1 : 0;
      events += r->row_ptr[i + 1] - r->row_ptr[i];
      for (uint64_t k = r->row_ptr[i]; k < r->row_ptr[i + 1]; ++k) {
        uint32_t j = r->col[k];
        int sl = (int)((t + r->delay[k]) % (uint64_t)D);
        float wk = r->w[k];
        int64_t q = (int64_t)llrint((double)wk * ldexp(1.0, r->S));
        r->ring[((size_t)sl * ch + chn) * n + j] += q;
        if (trip && i < ne && j < ne) { /* pre spike: depression by the post's last spike */
          float x = tab(r->mtab, r->mtab_n, (int64_t)ti - r->last[j]);
          float dlt = cfg->a_minus * x;
          float wn = wk - dlt;
          wn = wn < cfg->w_min ? cfg->w_min : (wn > cfg->w_max ? cfg->w_max : wn);
          r->w[k] = wn;
        }
        if (pair && i < ne && j < ne) {
          float p = tab(r->mtab, r->mtab_n, (int64_t)ti - r->post_s[j]);
          float x = r->post_v[j] * p;
          float dlt = cfg->a_plus * x;
          float wn = wk + dlt;
          wn = wn < cfg->w_min ? cfg->w_min : (wn > cfg->w_max ? cfg->w_max : wn);
          r->w[k] = wn;
        }
      }
    }
    /* ---- depression on E->E in-synapses of spiking excitatory targets, then post bump ---- */
    if (pair) {
      for (uint64_t s = 0; s < ns; ++s) {
        uint32_t j = spk[s];
        if (j >= ne) continue;
        for (uint64_t e = r->csc_ptr[j]; e < r->csc_ptr[j + 1]; ++e) {
          uint32_t i = r->csc_src[e];
          uint64_t k = r->csc_slot[e];
          float p = tab(r->ptab, r->ptab_n, (int64_t)ti - r->pre_s[i]);
          float x = r->pre_v[i] * p;
          float dlt = cfg->a_minus * x;
          float wn = r->w[k] - dlt;
          wn = wn < cfg->w_min ? cfg->w_min : (wn > cfg->w_max ? cfg->w_max : wn);
          r->w[k] = wn;
        }
        float p = tab(r->mtab, r->mtab_n, (int64_t)ti - r->post_s[j]);
        float x = r->post_v[j] * p;
        x = x * p;
        r->post_v[j] = x + 1.0f;
        r->post_s[j] = ti;
      }
    }
    /* ---- triplet: potentiation on E->E in-synapses of spiking excitatory targets ---- */
    if (trip) {
      for (uint64_t s = 0; s < ns; ++s) {
        uint32_t j = spk[s];
        if (j >= ne) continue;
        for (uint64_t e = r->csc_ptr[j]; e < r->csc_ptr[j + 1]; ++e) {
          uint32_t i = r->csc_src[e];
          uint64_t k = r->csc_slot[e];
          int32_t l = r->last[i];
          if (l == ti || l == NEVER_STEP) continue; /* same step: the pre spike's LTD only */
          float x1 = tab(r->ptab, r->ptab_n, (int64_t)ti - l);
          float x2 = tab(r->ytab, r->ytab_n, (int64_t)ti - r->post_s[j]);
          float dw = cfg->a_plus * x1;
          float y = cfg->a3_plus * x2;
          dw = dw + y;
          float wn = r->w[k] + dw;
          wn = wn < cfg->w_min ? cfg->w_min : (wn > cfg->w_max ? cfg->w_max : wn);
          r->w[k] = wn;
        }
      }
    }
  }
  free(spk);
  if (events_out) *events_out = events;
  return count;
}

void mir_read_state(mir_net* r, nsx_state_view* s) {
  size_t b = r->n * 4;
  if (s->v) memcpy(s->v, r->v, b);
  if (s->u) memcpy(s->u, r->u, b);
  if (s->m) memcpy(s->m, r->gm, b);
  if (s->h) memcpy(s->h, r->gh, b);
  if (s->n) memcpy(s->n, r->gn, b);
  if (s->syn0) memcpy(s->syn0, r->s0, b);
  if (s->syn1) memcpy(s->syn1, r->s1, b);
  if (s->syn2) memcpy(s->syn2, r->s2, b);
  if (s->drive) memcpy(s->drive, r->drive, b);
  if (s->refractory) memcpy(s->refractory, r->refr, b);
  if (s->last_spike_step) memcpy(s->last_spike_step, r->last, b);
  if (s->spike_count) memcpy(s->spike_count, r->count, b);
  if (s->pre_trace) memcpy(s->pre_trace, r->pre_v, b);
  if (s->pre_step) memcpy(s->pre_step, r->pre_s, b);
  if (s->post_trace) memcpy(s->post_trace, r->post_v, b);
  if (s->post_step) memcpy(s->post_step, r->post_s, b);
  if (s->ring) memcpy(s->ring, r->ring, (size_t)r->D * r->ch * r->n * 8);
  if (s->izh_c) memcpy(s->izh_c, r->c, b);
  if (s->izh_d) memcpy(s->izh_d, r->d, b);
}

void mir_write_state(mir_net* r, const nsx_state_view* s) {
  size_t b = r->n * 4;
  if (s->v) memcpy(r->v, s->v, b);
  if (s->u) memcpy(r->u, s->u, b);
  if (s->m) memcpy(r->gm, s->m, b);
  if (s->h) memcpy(r->gh, s->h, b);
  if (s->n) memcpy(r->gn, s->n, b);
  if (s->syn0) memcpy(r->s0, s->syn0, b);
  if (s->syn1) memcpy(r->s1, s->syn1, b);
  if (s->syn2) memcpy(r->s2, s->syn2, b);
  if (s->drive) memcpy(r->drive, s->drive, b);
  if (s->refractory) memcpy(r->refr, s->refractory, b);
  if (s->last_spike_step) memcpy(r->last, s->last_spike_step, b);
  if (s->spike_count) memcpy(r->count, s->spike_count, b);
  if (s->pre_trace) memcpy(r->pre_v, s->pre_trace, b);
  if (s->pre_step) memcpy(r->pre_s, s->pre_step, b);
  if (s->post_trace) memcpy(r->post_v, s->post_trace, b);
  if (s->post_step) memcpy(r->post_s, s->post_step, b);
  if (s->ring) memcpy(r->ring, s->ring, (size_t)r->D * r->ch * r->n * 8);
  if (s->izh_c) memcpy(r->c, s->izh_c, b);
  if (s->izh_d) memcpy(r->d, s->izh_d, b);
}
