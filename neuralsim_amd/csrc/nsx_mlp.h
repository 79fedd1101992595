// nsx_mlp.h -- the HH surrogate MLP 5 -> 64 -> 64 -> 4 (ReLU) with bf16 weights and FP32
// accumulation, evaluated and differentiated by one wavefront per chunk of 64 samples.
//
// All three layers run on gfx950 bf16 MFMA (v_mfma_f32_32x32x16_bf16 / v_mfma_f32_16x16x32_bf16)
// with activations rounded to bf16.  Layer 1 (K = 5) carries each normalised input as two bf16
// words, x = hi + lo (lo = bf16(x - hi), 16 significant bits), and its bias as two words against
// constant-1 inputs: K = 5 + 5 + 2 <= 16, one MFMA per 32x32 tile, so neither the inputs nor the
// bias are cut to 8 bits.  Layer 2's bias initialises the accumulators.  The output layer's
// weights are carried as two bf16 words, W3 = hi + lo: its rounding error lands directly on the
// outputs, and a single bf16 word there is the largest error of the whole surrogate step.  ReLU
// runs on the packed bf16 results as a signed 16-bit max against 0 (the sign bit of a bf16 is
// that of an int16, so relu-then-round and round-then-max agree).  Operand layouts:
//   32x32x16: lane l holds A[m = l%32][k = 8(l/32)+j], B[k = 8(l/32)+j][n = l%32] (j < 8);
//             C register i holds (m = 8(i/4) + 4(l/32) + i%4, n = l%32)
//   16x16x32: lane l holds A[m = l%16][k = 8(l/16)+j], B[k = 8(l/16)+j][n = l%16];
//             C register i holds (m = 4(l/16) + i, n = l%16)
//
// Inference (mlp_infer) keeps every activation in registers.  Samples are the n index of
// 32x32x16 tiles throughout (lane l's sample is n = l%32 of tile l/32), so a layer's C tile is
// the next layer's B operand as it stands: C registers 8q..8q+7 of the tile for output rows
// 32mt.. hold, in lane half l/32, the 16 rows 16(2mt+q) + 8(j/4) + 4(l/32) + j%4 (j < 8) -- one
// k-step of 16 rows in a fixed permuted order.  The next layer's A fragments store their K
// columns in that order (w2pf, w3pf), so the contraction is unchanged and no activation passes
// through LDS.  The layer-1 operand is each lane's own input row, moved to tile order by one swap
// of lane halves; the output layer stacks W3's hi words (rows 0-3) and lo words (rows 4-7) in one
// 32-row tile, whose hi and lo halves one more swap of lane halves brings to the sample's lane.
// Training (mlp_train_chunk) stages operand tiles K-contiguous in LDS ([m][k] for A, [n][k] for
// B), so a lane's eight bf16 are one 16-byte LDS read.  The weight operands are pre-arranged per
// lane in global memory (MlpImage) and read with coalesced 16-byte loads.
#pragma once
#include <stddef.h>

#include "nsx_device.h"

namespace nsx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMlpIn = 5, kMlpH = 64, kMlpOut = 4;
// FP32 master parameters: w1[64][5] b1[64] w2[64][64] b2[64] w3[4][64] b3[4]
constexpr int kOffW1 = 0, kOffB1 = 320, kOffW2 = 384, kOffB2 = 4480, kOffW3 = 4544,
              kOffB3 = 4800, kMlpParams = 4804;
constexpr int kLdA = 72;  // bf16 row length of 64-wide tiles (rows 16-byte aligned, padded)
constexpr int kLdO = 24;  // bf16 row length of 16-wide tiles

// Normalisation and training envelope, raw units: x_n = (x - in_mu) * in_isd, output in target
// units y = y_n * out_sd + out_mu; a state is inside the envelope when lo <= x <= hi.  The
// surrogate applies y * (|dV| + scale_eps), dV the frozen-rate step's change of V.
struct MlpNorm {
  float in_mu[kMlpIn], in_isd[kMlpIn];
  float out_mu[kMlpOut], out_sd[kMlpOut];
  float lo[kMlpIn], hi[kMlpIn];
  float scale_eps;
  float pad[3];
};
constexpr int kNormFloats = (int)(sizeof(MlpNorm) / 4);

// Layer-1 columns (K = 16): the inputs' high words, their low words, two constant-1 columns
// against the bias's two words, zero padding.
constexpr int kL1Lo = kMlpIn, kL1Bias = 2 * kMlpIn;

// Weights as the kernels read them (rebuilt from the FP32 masters after every optimiser step).
// The inference fragments (w1f, w2pf, w3pf) come first and are contiguous.
struct __attribute__((aligned(16))) MlpImage {
  float b2[kMlpH], b3[kMlpOut];  // FP32
  bf16x8 w1f[2][64];   // A fragments of [W1 | W1 | b1 hi | b1 lo | 0] [h1][16]: m-tile of 32
  bf16x8 w2pf[8][64];  // A fragments of W2 [h2][h1], tile mt*4+ks, K in accumulator order
  bf16x8 w3pf[4][64];  // A fragments of [W3 hi; W3 lo; 0] [32][h2], k-step ks, accumulator order
  bf16x8 w2f[8][64];   // A fragments of W2 [h2][h1]: tile mt*4+ks (m-tile of 32, k-step of 16)
  bf16x8 w3f[2][64];   // A fragments of W3 [o][h2] (hi words), o padded to 16: k-step of 32
  bf16x8 w3lf[2][64];  // the same for the lo words
  bf16x8 w2tf[8][64];  // A fragments of W2^T [h1][h2]: tile mt*4+ks
  bf16x8 w3tf[2][64];  // A fragments of W3^T [h2][o] (hi words), o padded to 16: m-tile of 32
  bf16x8 w3tlf[2][64]; // the same for the lo words
};
constexpr int kImageFrags = 38;  // fragment arrays of 64 lanes
static_assert(offsetof(MlpImage, w1f) % 16 == 0, "fragment alignment");

// Hidden row held at element j (< 8) of k-step ks's operand in lane half `half`: the row order of
// C registers 8(ks%2)..8(ks%2)+7 of a 32-row tile of rows 32(ks/2)..
__host__ __device__ constexpr uint32_t acc_row(uint32_t ks, uint32_t half, uint32_t j) {
  return 16u * ks + 8u * (j >> 2) + 4u * half + (j & 3u);
}

// The image's FP32 head and the normalisation, as staged into LDS by every block that evaluates
// the MLP: read with wave-uniform addresses, they would otherwise be vector loads with a wait each
// (the kernels' own global stores keep the compiler from proving them invariant for scalar loads).
struct __attribute__((aligned(16))) MlpScalars {
  float b2[kMlpH], b3[kMlpOut];
  MlpNorm norm;
  float b2p[kMlpH];  // b2 per (lane half, m-tile) in C-register order: [half*2 + mt][i]
};
constexpr int kScalarHead = kMlpH + kMlpOut;  // floats taken from the image
constexpr int kScalarFloats = kScalarHead + kNormFloats;
static_assert(offsetof(MlpImage, w1f) == kScalarHead * 4 &&
                  offsetof(MlpScalars, b2p) == kScalarFloats * 4 && kScalarFloats % 4 == 0,
              "the image starts with b2, b3; b2p follows the normalisation, 16-byte aligned");

// Cooperative copy by the whole block; the caller synchronises before use.
__device__ inline void stage_scalars(MlpScalars& dst, const MlpImage* img, const MlpNorm* norm) {
  float* d = reinterpret_cast<float*>(&dst);
  const float* a = reinterpret_cast<const float*>(img);
  const float* b = reinterpret_cast<const float*>(norm);
  for (int k = threadIdx.x; k < kScalarFloats + kMlpH; k += blockDim.x) {
    if (k < kScalarHead) {
      d[k] = a[k];
    } else if (k < kScalarFloats) {
      d[k] = b[k - kScalarHead];
    } else {  // b2p[(half*2 + mt)*16 + i] = b2[32mt + 8(i/4) + 4 half + i%4]
      const int q = k - kScalarFloats, i = q & 15, mt = (q >> 4) & 1, half = q >> 5;
      d[k] = a[32 * mt + 8 * (i >> 2) + 4 * half + (i & 3)];
    }
  }
}

// The inference fragments of the image (w1f .. w3pf, contiguous in MlpImage), staged into LDS by
// kernels that evaluate the MLP many times per block: a wavefront then reads each operand from
// LDS at its use instead of waiting on a global (L1 / L2) load before each MFMA.
struct __attribute__((aligned(16))) MlpInferW {
  bf16x8 w1f[2][64];
  bf16x8 w2pf[8][64];
  bf16x8 w3pf[4][64];
};
static_assert(offsetof(MlpImage, w2pf) - offsetof(MlpImage, w1f) == offsetof(MlpInferW, w2pf) &&
                  offsetof(MlpImage, w3pf) - offsetof(MlpImage, w1f) == offsetof(MlpInferW, w3pf) &&
                  offsetof(MlpImage, w2f) - offsetof(MlpImage, w1f) == sizeof(MlpInferW),
              "the inference fragments are contiguous in the image");

// Cooperative copy by the whole block; the caller synchronises before use.
__device__ inline void stage_infer_weights(MlpInferW& dst, const MlpImage* img) {
  typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
  u32x4* d = reinterpret_cast<u32x4*>(&dst);
  const u32x4* a = reinterpret_cast<const u32x4*>(img->w1f);
  for (uint32_t k = threadIdx.x; k < sizeof(MlpInferW) / 16; k += blockDim.x) d[k] = a[k];
}

__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline bf16x8 ld8(const __bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }

// acc += A B^T over a 32x32x16 tile; A from a register fragment, B[n][k] from LDS (row ldb).
__device__ inline void mma32(f32x16& acc, bf16x8 a, const __bf16* B, int ldb, uint32_t lane) {
  const bf16x8 b = ld8(B + (lane & 31) * ldb + (lane >> 5) * 8);
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
}
// Same with A[m][k] from LDS (row lda).
__device__ inline void mma32(f32x16& acc, const __bf16* A, int lda, const __bf16* B, int ldb,
                             uint32_t lane) {
  mma32(acc, ld8(A + (lane & 31) * lda + (lane >> 5) * 8), B, ldb, lane);
}
// acc += A B^T over a 16x16x32 tile.
__device__ inline void mma16(f32x4& acc, bf16x8 a, const __bf16* B, int ldb, uint32_t lane) {
  const bf16x8 b = ld8(B + (lane & 15) * ldb + (lane >> 4) * 8);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
}
__device__ inline void mma16(f32x4& acc, const __bf16* A, int lda, const __bf16* B, int ldb,
                             uint32_t lane) {
  mma16(acc, ld8(A + (lane & 15) * lda + (lane >> 4) * 8), B, ldb, lane);
}

// Row (m) of C register i of a 32x32 tile.
__device__ inline uint32_t c32_row(uint32_t i, uint32_t lane) {
  return 8u * (i >> 2) + 4u * (lane >> 5) + (i & 3u);
}

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef short i16x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// Two FP32 values rounded to bf16 (nearest even) in one packed conversion.
__device__ inline uint32_t pack_bf16x2(float a, float b) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, bf16x2));
}
// relu of two FP32 values as packed bf16: round first, then a signed 16-bit max against 0.
__device__ inline uint32_t relu_bf16x2(float a, float b) {
  const i16x2 q = __builtin_bit_cast(i16x2, pack_bf16x2(a, b));
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(q, (i16x2){0, 0}));
}
__device__ inline __bf16 bf16_half(uint32_t pair, uint32_t hi) {
  return __builtin_bit_cast(__bf16, (unsigned short)(hi ? pair >> 16 : pair & 0xffffu));
}

// relu of a 32x32 accumulator tile (m = h0 + row, n = s0 + column) as bf16 into an [s][h] tile
// (sh) and, when given, an [h][s] tile (hs).  Lane l < 32 holds rows 8g..8g+3 of column l and
// lane l + 32 rows 8g+4..8g+7; one swap of lane halves (v_permlane32_swap) between the register
// groups g = 2j and 2j + 1 gives lane l rows 16j..16j+7 and lane l + 32 rows 16j+8..16j+15, so
// each lane writes two 16-byte row pieces.  32 column-wise 8-byte stores with a 144-byte row
// pitch fall on 16 bank pairs (two lanes each); 16-byte stores of 16 lanes cover the 64 banks.
__device__ inline void store_relu(const f32x16& c, uint32_t h0, uint32_t s0, __bf16* sh,
                                  __bf16* hs, uint32_t lane) {
  typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
  const uint32_t s = s0 + (lane & 31);
  u32x2 r[4];
#pragma unroll
  for (uint32_t g = 0; g < 4; ++g) {
    r[g] = u32x2{relu_bf16x2(c[4 * g], c[4 * g + 1]), relu_bf16x2(c[4 * g + 2], c[4 * g + 3])};
    if (hs) {
#pragma unroll
      for (uint32_t q = 0; q < 4; ++q)
        hs[(h0 + c32_row(4 * g + q, lane)) * kLdA + s] = bf16_half(r[g][q >> 1], q & 1u);
    }
  }
#pragma unroll
  for (uint32_t j = 0; j < 2; ++j) {
    const auto x = __builtin_amdgcn_permlane32_swap(r[2 * j][0], r[2 * j + 1][0], false, false);
    const auto y = __builtin_amdgcn_permlane32_swap(r[2 * j][1], r[2 * j + 1][1], false, false);
    *reinterpret_cast<u32x4*>(sh + s * kLdA + h0 + 16 * j + 8 * (lane >> 5)) =
        u32x4{x[0], y[0], x[1], y[1]};
  }
}

// Layer 1 of a 64-sample tile: relu(W1 x + b1) as bf16 into act ([s][h1], row kLdA) and, when
// given, an [h1][s] tile (hs).  The lane's input row (x hi words, lo words, 1, 1, 0...) passes
// through elements 0..15 of act's row `lane`, which the outputs then overwrite.
template <class WT>
__device__ inline void layer1(const WT& W, const float xn[kMlpIn], __bf16* act, __bf16* hs,
                              uint32_t lane) {
  float hi[kMlpIn], lo[kMlpIn];
#pragma unroll
  for (int k = 0; k < kMlpIn; ++k) {
    hi[k] = (float)(__bf16)xn[k];
    lo[k] = xn[k] - hi[k];  // exact; rounded to bf16 when packed
  }
  typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
  const uint32_t one = 0x3f803f80u;  // two bf16 1.0
  const u32x4 r0 = {pack_bf16x2(hi[0], hi[1]), pack_bf16x2(hi[2], hi[3]),
                    pack_bf16x2(hi[4], lo[0]), pack_bf16x2(lo[1], lo[2])};
  const u32x4 r1 = {pack_bf16x2(lo[3], lo[4]), one, 0u, 0u};
  *reinterpret_cast<u32x4*>(act + lane * kLdA) = r0;
  *reinterpret_cast<u32x4*>(act + lane * kLdA + 8) = r1;
  wave_lds_sync();
  const bf16x8 x0 = ld8(act + (lane & 31) * kLdA + (lane >> 5) * 8);         // samples 0..31
  const bf16x8 x1 = ld8(act + (32 + (lane & 31)) * kLdA + (lane >> 5) * 8);  // samples 32..63
  wave_lds_sync();  // every lane holds its operands before the rows are overwritten
#pragma unroll
  for (uint32_t mt = 0; mt < 2; ++mt) {
    const bf16x8 a = W.w1f[mt][lane];
    const f32x16 c0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, x0, f32x16{}, 0, 0, 0);
    const f32x16 c1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, x1, f32x16{}, 0, 0, 0);
    store_relu(c0, mt * 32, 0, act, hs, lane);
    store_relu(c1, mt * 32, 32, act, hs, lane);
  }
}

// Layer-2 pre-activations: acc[mt*2+nt] (h2 tile mt, sample tile nt) = b2 + W2 a1^T, a1 as
// [s][h1]; the bias initialises the accumulators.
template <class WT>
__device__ inline void layer2(const WT& W, const float* b2, const __bf16* a1_sh,
                              f32x16 acc[4], uint32_t lane) {
#pragma unroll
  for (uint32_t mt = 0; mt < 2; ++mt) {
    f32x16 b;
#pragma unroll
    for (uint32_t i = 0; i < 16; ++i) b[i] = b2[mt * 32 + c32_row(i, lane)];
    acc[mt * 2] = b;
    acc[mt * 2 + 1] = b;
  }
#pragma unroll
  for (uint32_t ks = 0; ks < 4; ++ks)
#pragma unroll
    for (uint32_t mt = 0; mt < 2; ++mt) {
      const bf16x8 a = W.w2f[mt * 4 + ks][lane];
#pragma unroll
      for (uint32_t nt = 0; nt < 2; ++nt)
        mma32(acc[mt * 2 + nt], a, a1_sh + nt * 32 * kLdA + ks * 16, kLdA, lane);
    }
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// relu of a 32x32 accumulator tile as bf16, packed as the next layer's two k-step operands:
// op[q] holds C registers 8q..8q+7 (rows acc_row(2mt + q, lane half, j) of the tile's m-tile mt).
__device__ inline void relu_operands(const f32x16& c, bf16x8 op[2]) {
#pragma unroll
  for (uint32_t q = 0; q < 2; ++q) {
    const u32x4 r = {relu_bf16x2(c[8 * q], c[8 * q + 1]), relu_bf16x2(c[8 * q + 2], c[8 * q + 3]),
                     relu_bf16x2(c[8 * q + 4], c[8 * q + 5]), relu_bf16x2(c[8 * q + 6], c[8 * q + 7])};
    op[q] = __builtin_bit_cast(bf16x8, r);
  }
}

// Normalised outputs of the calling lane's sample (normalised inputs xn), with every activation
// in registers (layouts in the header comment).  Every lane of the wavefront must call it (the
// MFMAs use all 64 lanes); no LDS is touched except the weights W (LDS or global) and S.
template <class WT>
__device__ inline void mlp_infer(const WT& W, const MlpScalars& S, const float xn[kMlpIn],
                                 float y[kMlpOut], uint32_t lane) {
  const uint32_t half = lane >> 5;
  float hi[kMlpIn], lo[kMlpIn];
#pragma unroll
  for (int k = 0; k < kMlpIn; ++k) {
    hi[k] = (float)(__bf16)xn[k];
    lo[k] = xn[k] - hi[k];  // exact; rounded to bf16 when packed
  }
  // the lane's input row: columns 0-7 (hi 0-4, lo 0-2) and 8-15 (lo 3-4, 1, 1, 0...)
  const uint32_t one = 0x3f803f80u;  // two bf16 1.0
  const u32x4 r0 = {pack_bf16x2(hi[0], hi[1]), pack_bf16x2(hi[2], hi[3]),
                    pack_bf16x2(hi[4], lo[0]), pack_bf16x2(lo[1], lo[2])};
  const u32x4 r1 = {pack_bf16x2(lo[3], lo[4]), one, 0u, 0u};
  // one swap of lane halves: x[0] = operand of samples 0..31 (lanes 32..63 carry columns 8-15
  // of samples 0..31), x[1] = operand of samples 32..63
  u32x4 xb[2];
#pragma unroll
  for (uint32_t q = 0; q < 4; ++q) {
    const auto sw = __builtin_amdgcn_permlane32_swap(r0[q], r1[q], false, false);
    xb[0][q] = sw[0];
    xb[1][q] = sw[1];
  }
  f32x16 out[2];
#pragma unroll
  for (uint32_t nt = 0; nt < 2; ++nt) {  // sample tile: samples 32nt .. 32nt+31
    // one tile at a time: the tiles' chains need not overlap (other wavefronts fill the MFMA
    // pipe), and keeping them apart holds the live registers to one tile's operands
    if (nt) __builtin_amdgcn_sched_barrier(0);
    const bf16x8 xo = __builtin_bit_cast(bf16x8, xb[nt]);
    bf16x8 a1[4];  // layer-1 activations, k-steps of layer 2
#pragma unroll
    for (uint32_t mt = 0; mt < 2; ++mt)
      relu_operands(__builtin_amdgcn_mfma_f32_32x32x16_bf16(W.w1f[mt][lane], xo, f32x16{}, 0, 0, 0),
                    a1 + 2 * mt);
    bf16x8 a2[4];  // layer-2 activations, k-steps of layer 3
#pragma unroll
    for (uint32_t mt = 0; mt < 2; ++mt) {
      f32x16 acc = *reinterpret_cast<const f32x16*>(S.b2p + (half * 2 + mt) * 16);
#pragma unroll
      for (uint32_t ks = 0; ks < 4; ++ks)
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W.w2pf[mt * 4 + ks][lane], a1[ks], acc, 0, 0,
                                                      0);
      relu_operands(acc, a2 + 2 * mt);
    }
    f32x16 o = f32x16{};
#pragma unroll
    for (uint32_t ks = 0; ks < 4; ++ks)
      o = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W.w3pf[ks][lane], a2[ks], o, 0, 0, 0);
    out[nt] = o;
  }
  // C registers 0..3: rows 0..3 (hi words' products) in lanes 0..31, rows 4..7 (lo words') in
  // lanes 32..63, for sample 32nt + l%32; one swap gives lane l both parts of sample l.  Vector
  // elements are copied to scalars before __builtin_bit_cast: this compiler's bit_cast of a
  // vector element expression reads element 0.
#pragma unroll
  for (uint32_t q = 0; q < 4; ++q) {
    const float e0 = out[0][q], e1 = out[1][q];
    const auto sw = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(uint32_t, e0),
                                                     __builtin_bit_cast(uint32_t, e1), false,
                                                     false);
    const uint32_t h = sw[0], l = sw[1];
    float v = __builtin_bit_cast(float, h) + __builtin_bit_cast(float, l);
    y[q] = v + S.b3[q];
  }
}

// LDS of one training wavefront (forward and backward of a 64-sample chunk); 62,976 bytes.
struct __attribute__((aligned(16))) TrainLds {
  float x[64 * kMlpIn];  // normalised inputs [s][k]
  float d3[64 * 4];      // normalised targets, then FP32 output gradients [s][o]
  union {
    struct {
      __bf16 a1_sh[64 * kLdA];  // layer-1 activations [s][h1]
      __bf16 a2_sh[64 * kLdA];  // layer-2 activations [s][h2]
    } f;
    float dz1[64 * 65];  // dL/dz1 [h1][s] (FP32, odd row length: conflict-free column walks)
  } u;
  __bf16 a1_hs[64 * kLdA];   // [h1][s]
  __bf16 a2_hs[64 * kLdA];   // [h2][s]
  __bf16 dz2_hs[64 * kLdA];  // dL/dz2 [h2][s]
  __bf16 dz2_sh[64 * kLdA];  // dL/dz2 [s][h2]
  __bf16 dy_os[16 * kLdA];   // dL/dy [o][s], rows 4..15 zero
  __bf16 dy_so[64 * kLdO];   // dL/dy [s][o], columns 4..15 zero
};
static_assert(offsetof(TrainLds, u) % 16 == 0 && offsetof(TrainLds, a1_hs) % 16 == 0 &&
                  offsetof(TrainLds, dy_os) % 16 == 0 && offsetof(TrainLds, dy_so) % 16 == 0,
              "LDS tile alignment");

// Gradient accumulators of one wavefront (MFMA C layouts for W2 and W3).
struct GradAcc {
  f32x16 w2[4];        // tile mt*2+nt: rows h2 (mt), columns h1 (nt)
  f32x4 w3[4];         // tile nt: rows o (lanes 0..15), columns h2 = nt*16 + lane%16
  float w1[kMlpIn];    // lane h1
  float b1, b2, b3;    // lane h (b3: lanes 0..3)
  float loss;          // sum of squared normalised errors of this lane's outputs
};

__device__ inline void grad_zero(GradAcc& G) {
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    G.w2[t] = f32x16{};
    G.w3[t] = f32x4{};
  }
#pragma unroll
  for (int k = 0; k < kMlpIn; ++k) G.w1[k] = 0.0f;
  G.b1 = G.b2 = G.b3 = G.loss = 0.0f;
}

// Forward and backward of one chunk.  The caller has written L.x (normalised inputs) and L.d3
// (normalised targets) for its lane's sample and passes the same inputs in xn.  Loss is
// gscale * sum (y - target)^2, so dL/dy = 2 gscale (y - target).  Backward treats the bf16
// rounding of activations as the identity.
__device__ inline void mlp_train_chunk(const MlpImage& W, const MlpScalars& S, TrainLds& L,
                                       GradAcc& G, const float xn[kMlpIn], float gscale,
                                       uint32_t lane) {
  // ---- forward ----
  layer1(W, xn, L.u.f.a1_sh, L.a1_hs, lane);
  wave_lds_sync();
  f32x16 acc[4];
  layer2(W, S.b2, L.u.f.a1_sh, acc, lane);
#pragma unroll
  for (uint32_t mt = 0; mt < 2; ++mt)
#pragma unroll
    for (uint32_t nt = 0; nt < 2; ++nt)
      store_relu(acc[mt * 2 + nt], mt * 32, nt * 32, L.u.f.a2_sh, L.a2_hs, lane);
  wave_lds_sync();
  {
    const bf16x8 w30 = W.w3f[0][lane], w31 = W.w3f[1][lane];
    const bf16x8 w30l = W.w3lf[0][lane], w31l = W.w3lf[1][lane];
#pragma unroll
    for (uint32_t nt = 0; nt < 4; ++nt) {
      f32x4 o = f32x4{};
      mma16(o, w30l, L.u.f.a2_sh + nt * 16 * kLdA, kLdA, lane);
      mma16(o, w31l, L.u.f.a2_sh + nt * 16 * kLdA + 32, kLdA, lane);
      mma16(o, w30, L.u.f.a2_sh + nt * 16 * kLdA, kLdA, lane);
      mma16(o, w31, L.u.f.a2_sh + nt * 16 * kLdA + 32, kLdA, lane);
      const uint32_t s = nt * 16 + (lane & 15);
#pragma unroll
      for (uint32_t q = 0; q < 4; ++q) {
        const uint32_t row = 4 * (lane >> 4) + q;  // output index; rows >= 4 are padding
        float d = 0.0f;
        if (row < 4) {
          const float e = (o[q] + S.b3[row]) - L.d3[s * 4 + row];
          G.loss += e * e;
          d = 2.0f * gscale * e;
          L.d3[s * 4 + row] = d;
        }
        const __bf16 db = (__bf16)d;
        L.dy_os[row * kLdA + s] = db;
        L.dy_so[s * kLdO + row] = db;
      }
    }
  }
  wave_lds_sync();
  // ---- backward ----
  // dW3[o][h2] += dy[o][s] a2[h2][s]^T ; db3[o] += sum_s dy[o][s]
#pragma unroll
  for (uint32_t nt = 0; nt < 4; ++nt)
#pragma unroll
    for (uint32_t ks = 0; ks < 2; ++ks)
      mma16(G.w3[nt], L.dy_os + ks * 32, kLdA, L.a2_hs + nt * 16 * kLdA + ks * 32, kLdA, lane);
  if (lane < 4) {
    float sb = 0.0f;
    for (uint32_t s = 0; s < 64; ++s) sb += L.d3[s * 4 + lane];
    G.b3 += sb;
  }
  // dz2[h2][s] = (W3^T dy)[h2][s] where a2 > 0
#pragma unroll
  for (uint32_t mt = 0; mt < 2; ++mt) {
    const bf16x8 a = W.w3tf[mt][lane], al = W.w3tlf[mt][lane];
#pragma unroll
    for (uint32_t nt = 0; nt < 2; ++nt) {
      acc[mt * 2 + nt] = f32x16{};
      mma32(acc[mt * 2 + nt], al, L.dy_so + nt * 32 * kLdO, kLdO, lane);
      mma32(acc[mt * 2 + nt], a, L.dy_so + nt * 32 * kLdO, kLdO, lane);
    }
  }
#pragma unroll
  for (uint32_t mt = 0; mt < 2; ++mt)
#pragma unroll
    for (uint32_t nt = 0; nt < 2; ++nt) {
      const uint32_t s = nt * 32 + (lane & 31);
#pragma unroll
      for (uint32_t i = 0; i < 16; ++i) {
        const uint32_t h = mt * 32 + c32_row(i, lane);
        const bool on = (float)L.a2_hs[h * kLdA + s] > 0.0f;
        const __bf16 d = (__bf16)(on ? acc[mt * 2 + nt][i] : 0.0f);
        L.dz2_hs[h * kLdA + s] = d;
        L.dz2_sh[s * kLdA + h] = d;
      }
    }
  wave_lds_sync();
  {
    float sb = 0.0f;
    for (uint32_t s = 0; s < 64; ++s) sb += (float)L.dz2_hs[lane * kLdA + s];
    G.b2 += sb;
  }
  // dW2[h2][h1] += dz2[h2][s] a1[h1][s]^T
#pragma unroll
  for (uint32_t ks = 0; ks < 4; ++ks)
#pragma unroll
    for (uint32_t mt = 0; mt < 2; ++mt)
#pragma unroll
      for (uint32_t nt = 0; nt < 2; ++nt)
        mma32(G.w2[mt * 2 + nt], L.dz2_hs + mt * 32 * kLdA + ks * 16, kLdA,
              L.a1_hs + nt * 32 * kLdA + ks * 16, kLdA, lane);
  // dz1[h1][s] = (W2^T dz2)[h1][s] where a1 > 0 (FP32, into the dead a1_sh / a2_sh storage)
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = f32x16{};
#pragma unroll
  for (uint32_t ks = 0; ks < 4; ++ks)
#pragma unroll
    for (uint32_t mt = 0; mt < 2; ++mt) {
      const bf16x8 a = W.w2tf[mt * 4 + ks][lane];
#pragma unroll
      for (uint32_t nt = 0; nt < 2; ++nt)
        mma32(acc[mt * 2 + nt], a, L.dz2_sh + nt * 32 * kLdA + ks * 16, kLdA, lane);
    }
#pragma unroll
  for (uint32_t mt = 0; mt < 2; ++mt)
#pragma unroll
    for (uint32_t nt = 0; nt < 2; ++nt) {
      const uint32_t s = nt * 32 + (lane & 31);
#pragma unroll
      for (uint32_t i = 0; i < 16; ++i) {
        const uint32_t h = mt * 32 + c32_row(i, lane);
        const bool on = (float)L.a1_hs[h * kLdA + s] > 0.0f;
        L.u.dz1[h * 65 + s] = on ? acc[mt * 2 + nt][i] : 0.0f;
      }
    }
  wave_lds_sync();
  // dW1[h1][k] += dz1[h1][s] x[s][k] ; db1[h1] += sum_s dz1[h1][s]   (lane h1, FP32)
  {
    float sb = 0.0f;
    for (uint32_t s = 0; s < 64; ++s) {
      const float d = L.u.dz1[lane * 65 + s];
      sb += d;
#pragma unroll
      for (int k = 0; k < kMlpIn; ++k) G.w1[k] = __builtin_fmaf(d, L.x[s * kMlpIn + k], G.w1[k]);
    }
    G.b1 += sb;
  }
  wave_lds_sync();  // the next chunk overwrites every buffer
}

// Scatter a wavefront's accumulated gradient into master order (kMlpParams floats at out).
__device__ inline void grad_store(const GradAcc& G, float* out, uint32_t lane) {
#pragma unroll
  for (int k = 0; k < kMlpIn; ++k) out[kOffW1 + lane * kMlpIn + k] = G.w1[k];
  out[kOffB1 + lane] = G.b1;
  out[kOffB2 + lane] = G.b2;
#pragma unroll
  for (uint32_t mt = 0; mt < 2; ++mt)
#pragma unroll
    for (uint32_t nt = 0; nt < 2; ++nt)
#pragma unroll
      for (uint32_t i = 0; i < 16; ++i) {
        const uint32_t h2 = mt * 32 + c32_row(i, lane), h1 = nt * 32 + (lane & 31);
        out[kOffW2 + h2 * 64 + h1] = G.w2[mt * 2 + nt][i];
      }
  if (lane < 16) {
#pragma unroll
    for (uint32_t nt = 0; nt < 4; ++nt)
#pragma unroll
      for (uint32_t q = 0; q < 4; ++q) out[kOffW3 + q * 64 + nt * 16 + lane] = G.w3[nt][q];
  }
  if (lane < 4) out[kOffB3 + lane] = G.b3;
}

}  // namespace nsx
