// nsx_hh.h -- Hodgkin-Huxley Euler sub-step shared by the plain network step, the surrogate's
// training-data generator and its per-neuron fallback.  Operation order is fixed (the library
// is compiled with -ffp-contract=off) so the CPU restatement matches bit for bit.
#pragma once
#include "nsx_device.h"
#include "nsx_internal.h"

namespace nsx {

// alpha = scale * x / (1 - exp(-x/10)); near x = 0 the limit k * (1 + x/20)
// (alpha_m, alpha_n of hodgkin_huxley.rs:20-42 have a removable singularity there).
__device__ inline float hh_ratio(float x, float scale, float k) {
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

// Opening and closing rates of the three gates at membrane potential v (hodgkin_huxley.rs:20-42).
struct HHRates {
  float am, bm, ah, bh, an, bn;
};

__device__ inline HHRates hh_rates(float vo) {
  HHRates r;
  r.am = hh_ratio(vo + 40.0f, 0.1f, 1.0f);
  r.bm = 4.0f * expf_((vo + 65.0f) * (-1.0f / 18.0f));
  r.ah = 0.07f * expf_((vo + 65.0f) * -0.05f);
  float eh = expf_((vo + 35.0f) * -0.1f);
  eh = 1.0f + eh;
  r.bh = 1.0f / eh;
  r.an = hh_ratio(vo + 55.0f, 0.01f, 0.1f);
  r.bn = 0.125f * expf_((vo + 65.0f) * -0.0125f);
  return r;
}

// The rest of one Euler sub-step given the rates: gates by Euler, then the currents with the new
// gates and the old V, input current I held.  Standard mode integrates the documented equation
// C dV/dt = g_Na m^3 h (E_Na-V) + g_K n^4 (E_K-V) + g_L (E_L-V) + I (hodgkin_huxley.rs:6-9);
// literal mode subtracts those currents as the code does (hodgkin_huxley.rs:78-84).
__device__ inline void hh_apply(float& v, float& m, float& h, float& n, float I,
                                const HHRates& r, const HHConst& c) {
  const float vo = v;
  float t1, t2;
  t1 = r.am * (1.0f - m); t2 = r.bm * m; t1 = t1 - t2; t1 = t1 * c.dts; m = m + t1;
  t1 = r.ah * (1.0f - h); t2 = r.bh * h; t1 = t1 - t2; t1 = t1 * c.dts; h = h + t1;
  t1 = r.an * (1.0f - n); t2 = r.bn * n; t1 = t1 - t2; t1 = t1 * c.dts; n = n + t1;
  float m3 = m * m;
  m3 = m3 * m;
  float ina = c.g_na * m3;
  ina = ina * h;
  ina = ina * (c.e_na - vo);
  float n2 = n * n;
  float n4 = n2 * n2;
  float ik = c.g_k * n4;
  ik = ik * (c.e_k - vo);
  float il = c.g_l * (c.e_l - vo);
  float acc;
  if (c.literal) {
    acc = I - ina;
    acc = acc - ik;
    acc = acc - il;
  } else {
    acc = I + ina;
    acc = acc + ik;
    acc = acc + il;
  }
  acc = acc * c.kv;
  v = vo + acc;
}

// One Euler sub-step (HodgkinHuxleyNeuron::step order: gates at the old V, then V).
__device__ inline void hh_substep(float& v, float& m, float& h, float& n, float I,
                                  const HHConst& c) {
  const HHRates r = hh_rates(v);
  hh_apply(v, m, h, n, I, r, c);
}

// The surrogate's physics part: `substeps` sub-steps whose rates are frozen at the step-start V
// (one rate evaluation instead of one per sub-step).  Equal to hh_substep when substeps == 1;
// the learned correction supplies what the rates' change within the step contributes.
__device__ inline void hh_frozen(float& v, float& m, float& h, float& n, float I,
                                 const HHConst& c, uint32_t substeps) {
  const HHRates r = hh_rates(v);
  for (uint32_t k = 0; k < substeps; ++k) hh_apply(v, m, h, n, I, r, c);
}

// Conductances (AMPA, GABA-A, NMDA) and OU drive value of an HH neuron, as loaded from and
// stored to s0, s1, s2, drv.
struct HHSyn {
  float ga, gg, gn, drv;
};

template <int DRIVE>
__device__ inline HHSyn hh_syn_load(const DevPtrs& s, uint32_t i) {
  HHSyn y{s.s0[i], s.s1[i], s.s2[i], 0.0f};
  if (DRIVE == NSX_DRIVE_OU) y.drv = s.drv[i];
  return y;
}

template <int DRIVE>
__device__ inline void hh_syn_store(const DevPtrs& s, uint32_t i, const HHSyn& y) {
  s.s0[i] = y.ga;
  s.s1[i] = y.gg;
  s.s2[i] = y.gn;
  if (DRIVE == NSX_DRIVE_OU) s.drv[i] = y.drv;
}

// Synaptic conductances and external drive of a neuron for one step (y updated in place);
// returns the input current held over the step's sub-steps.  AMPA / NMDA open on excitatory
// arrivals a0 (NMDA through the Mg block at the step-start V, synapse/mod.rs:41-47), GABA-A on
// inhibitory arrivals |a1|; the currents use the step-start V.  OU drive: the current value is
// used and the process advanced by its exact one-step transition with the normal draw z.
// lfp receives |I_AMPA| + |I_NMDA| + |I_GABA| (the neuron's term of the field-potential sum).
template <int DRIVE>
__device__ inline float hh_input(const StepParams& p, HHSyn& y, float v, float a0, float a1,
                                 float z, float& lfp) {
  float ga = y.ga * p.dec_e;
  float xa = a0 * p.gmax_ampa;
  ga = ga + xa;
  float gn = y.gn * p.dec_n;
  if (p.use_nmda) {
    float e = expf_(-0.062f * v);
    float bl = p.mgc * e;
    bl = 1.0f + bl;
    float xn = a0 * p.gmax_nmda;
    xn = xn / bl;
    gn = gn + xn;
  }
  float gg = y.gg * p.dec_i;
  float xg = fabsf(a1) * p.gmax_gaba;
  gg = gg + xg;
  float de = p.e_exc - v, di = p.e_inh - v;
  float isyn = ga * de;
  float tn = gn * de;
  float tg = gg * di;
  lfp = fabsf(isyn);
  lfp = lfp + fabsf(tn);
  lfp = lfp + fabsf(tg);
  isyn = isyn + tn;
  isyn = isyn + tg;
  float iext = 0.0f;
  if (DRIVE == NSX_DRIVE_OU) {
    iext = y.drv;
    float o = iext * p.ou_k1;
    o = o + p.ou_k2;
    float q = p.ou_k3 * z;
    y.drv = o + q;
  } else if (DRIVE == NSX_DRIVE_GAUSSIAN) {
    float q = p.drive_sigma * z;
    iext = p.drive_mean + q;
  }
  y.ga = ga;
  y.gg = gg;
  y.gn = gn;
  return iext + isyn;
}

}  // namespace nsx
