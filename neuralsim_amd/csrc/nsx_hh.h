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

// One Euler sub-step with input current I held (HodgkinHuxleyNeuron::step order: gates by Euler
// at the old V, then currents with the new gates and the old V).  Standard mode integrates the
// documented equation C dV/dt = g_Na m^3 h (E_Na-V) + g_K n^4 (E_K-V) + g_L (E_L-V) + I
// (hodgkin_huxley.rs:6-9); literal mode subtracts those currents as the code does.
__device__ inline void hh_substep(float& v, float& m, float& h, float& n, float I,
                                  const HHConst& c) {
  const float vo = v;
  float am = hh_ratio(vo + 40.0f, 0.1f, 1.0f);
  float bm = 4.0f * expf_((vo + 65.0f) * (-1.0f / 18.0f));
  float ah = 0.07f * expf_((vo + 65.0f) * -0.05f);
  float eh = expf_((vo + 35.0f) * -0.1f);
  eh = 1.0f + eh;
  float bh = 1.0f / eh;
  float an = hh_ratio(vo + 55.0f, 0.01f, 0.1f);
  float bn = 0.125f * expf_((vo + 65.0f) * -0.0125f);
  float t1, t2;
  t1 = am * (1.0f - m); t2 = bm * m; t1 = t1 - t2; t1 = t1 * c.dts; m = m + t1;
  t1 = ah * (1.0f - h); t2 = bh * h; t1 = t1 - t2; t1 = t1 * c.dts; h = h + t1;
  t1 = an * (1.0f - n); t2 = bn * n; t1 = t1 - t2; t1 = t1 * c.dts; n = n + t1;
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

}  // namespace nsx
