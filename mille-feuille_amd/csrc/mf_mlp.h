// mf_mlp.h -- the tile-precision MLP (8 -> 32 -> 32 -> 4, ReLU) on the
// device: per-tile feature layout and the MFMA evaluation used both by
// mf_mlp_forward and inside the solver kernels.
//
// Evaluation is transposed so that tiles are the N dimension of
// v_mfma_f32_16x16x16f16: H^T = W . F^T.  With the 16x16x16 lane layout
//   A[i][k]: lane l holds i = l%16, k = 4*(l/16) + v      (v = 0..3)
//   B[k][j]: lane l holds k = 4*(l/16) + v, j = l%16
//   D[i][j]: lane l holds i = 4*(l/16) + v, j = l%16
// the D block of one layer (hidden units x 16 tiles) is, after ReLU and
// rounding to FP16, exactly the B operand of the next layer for that block
// of 16 hidden units, so a batch of 16 tiles goes through the network in 8
// MFMAs (layer 1: 2, layer 2: 4, layer 3: 2) with no data movement between
// lanes.  Weights and biases are FP16; products accumulate in FP32.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mf {

constexpr int kMlpIn = MF_MLP_IN, kMlpH1 = MF_MLP_H1, kMlpH2 = MF_MLP_H2, kMlpOut = MF_MLP_OUT;
constexpr int kMlpParams = MF_MLP_NPARAM;
// Offsets of the packed parameter vector w1[32][8], b1[32], w2[32][32],
// b2[32], w3[4][32], b3[4].
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpH1 * kMlpIn;
constexpr int kOffW2 = kOffB1 + kMlpH1;
constexpr int kOffB2 = kOffW2 + kMlpH2 * kMlpH1;
constexpr int kOffW3 = kOffB2 + kMlpH2;
constexpr int kOffB3 = kOffW3 + kMlpOut * kMlpH2;

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// Parameters pre-arranged in the MFMA lane layout: a[blk][lane] is the A
// operand of lane `lane` for block blk (0-1: W1 rows 16m.., 2-5: W2 block
// (m2, kb) at 2 + 2*m2 + kb, 6-7: W3 columns 16kb..), c1/c2/c3 the bias C
// operands.  Built on the host from the FP32 parameter vector.
struct MlpDevice {
  half4_t a[8][64];
  float4_t c1[2][64];
  float4_t c2[2][64];
  float4_t c3[64];
};

__device__ __forceinline__ half4_t relu_half(float4_t v) {
  half4_t h;
  h.x = (_Float16)fmaxf(v.x, 0.0f);
  h.y = (_Float16)fmaxf(v.y, 0.0f);
  h.z = (_Float16)fmaxf(v.z, 0.0f);
  h.w = (_Float16)fmaxf(v.w, 0.0f);
  return h;
}

// One batch of 16 tiles.  f: this lane's B operand of layer 1 (features
// 4*(lane/16)..+3 of tile lane%16; zero in lanes 32-63).  Returns the D block
// of layer 3: lanes 0-15 hold the 4 logits of tile lane%16.  The operands are
// read from `w` (L1-resident after the first batch) right where they are
// used (mf_mlp_forward; the solvers hold them in registers, mlp_eval16r).
__device__ __forceinline__ float4_t mlp_eval16(const MlpDevice *__restrict__ w, int lane, half4_t f) {
  // The scheduling barriers keep each stage's operand loads next to their
  // MFMAs (hoisted together they would need ~36 more registers at once).
  const float4_t h1a = __builtin_amdgcn_mfma_f32_16x16x16f16(w->a[0][lane], f, w->c1[0][lane], 0, 0, 0);
  const float4_t h1b = __builtin_amdgcn_mfma_f32_16x16x16f16(w->a[1][lane], f, w->c1[1][lane], 0, 0, 0);
  const half4_t g0 = relu_half(h1a), g1 = relu_half(h1b);
  __builtin_amdgcn_sched_barrier(0);
  float4_t h2a = __builtin_amdgcn_mfma_f32_16x16x16f16(w->a[2][lane], g0, w->c2[0][lane], 0, 0, 0);
  h2a = __builtin_amdgcn_mfma_f32_16x16x16f16(w->a[3][lane], g1, h2a, 0, 0, 0);
  __builtin_amdgcn_sched_barrier(0);
  float4_t h2b = __builtin_amdgcn_mfma_f32_16x16x16f16(w->a[4][lane], g0, w->c2[1][lane], 0, 0, 0);
  h2b = __builtin_amdgcn_mfma_f32_16x16x16f16(w->a[5][lane], g1, h2b, 0, 0, 0);
  const half4_t q0 = relu_half(h2a), q1 = relu_half(h2b);
  __builtin_amdgcn_sched_barrier(0);
  float4_t z = __builtin_amdgcn_mfma_f32_16x16x16f16(w->a[6][lane], q0, w->c3[lane], 0, 0, 0);
  z = __builtin_amdgcn_mfma_f32_16x16x16f16(w->a[7][lane], q1, z, 0, 0, 0);
  return z;
}

// Index of the largest logit (first one on ties): the predicted level.
__device__ __forceinline__ int mlp_argmax(float4_t z) {
  int c = 0;
  float m = z.x;
  if (z.y > m) { m = z.y; c = 1; }
  if (z.z > m) { m = z.z; c = 2; }
  if (z.w > m) { c = 3; }
  return c;
}

// The MLP's operands of one lane, loaded once into registers for a run of
// batches (the solvers' in-kernel inference: about 36 VGPRs, live only while
// the wave evaluates its tiles, outside the SpMV loops).
struct MlpRegs {
  half4_t a[8];
  float4_t c1[2], c2[2], c3;
};

__device__ __forceinline__ void mlp_load(const MlpDevice *__restrict__ w, int lane, MlpRegs &r) {
#pragma unroll
  for (int k = 0; k < 8; ++k) r.a[k] = w->a[k][lane];
  r.c1[0] = w->c1[0][lane];
  r.c1[1] = w->c1[1][lane];
  r.c2[0] = w->c2[0][lane];
  r.c2[1] = w->c2[1][lane];
  r.c3 = w->c3[lane];
}

// mlp_eval16 with the operands in registers: the eight MFMAs wait only on
// each other, not on operand loads.
__device__ __forceinline__ float4_t mlp_eval16r(const MlpRegs &w, half4_t f) {
  const float4_t h1a = __builtin_amdgcn_mfma_f32_16x16x16f16(w.a[0], f, w.c1[0], 0, 0, 0);
  const float4_t h1b = __builtin_amdgcn_mfma_f32_16x16x16f16(w.a[1], f, w.c1[1], 0, 0, 0);
  const half4_t g0 = relu_half(h1a), g1 = relu_half(h1b);
  float4_t h2a = __builtin_amdgcn_mfma_f32_16x16x16f16(w.a[2], g0, w.c2[0], 0, 0, 0);
  h2a = __builtin_amdgcn_mfma_f32_16x16x16f16(w.a[3], g1, h2a, 0, 0, 0);
  float4_t h2b = __builtin_amdgcn_mfma_f32_16x16x16f16(w.a[4], g0, w.c2[1], 0, 0, 0);
  h2b = __builtin_amdgcn_mfma_f32_16x16x16f16(w.a[5], g1, h2b, 0, 0, 0);
  const half4_t q0 = relu_half(h2a), q1 = relu_half(h2b);
  float4_t z = __builtin_amdgcn_mfma_f32_16x16x16f16(w.a[6], q0, w.c3, 0, 0, 0);
  z = __builtin_amdgcn_mfma_f32_16x16x16f16(w.a[7], q1, z, 0, 0, 0);
  return z;
}

// In-kernel inference context: the MLP, the per-tile static features (8
// FP16 per tile; slots 6 and 7 are filled here) and the per-segment squared
// norms of the vector the SpMV multiplies.
struct InferCtx {
  const MlpDevice *w;
  const uint16_t *feat16;  // tilenum x 8
  const float *seg2;       // tilem: ||v_segment||^2
  uint8_t *prec_out;       // tilenum: levels chosen (kept for later iterations)
  float seg_scale;         // sqrt(nseg) / ||v||
  float progress;          // feature 7: log10(||r|| / ||b||) / 16
};

__device__ __forceinline__ float feature_local_norm(float seg2, float seg_scale) {
  // log2(||v_seg|| * sqrt(nseg) / ||v|| + 2^-40) / 40
  return __log2f(sqrtf(seg2) * seg_scale + 9.094947e-13f) * (1.0f / 40.0f);
}

// B operand of layer 1 for tile t (valid: the batch slot holds a tile).
__device__ __forceinline__ half4_t infer_features(const InferCtx &c, int t, int tcol, int lane,
                                                  bool valid) {
  half4_t f = {0, 0, 0, 0};
  if (!valid || lane >= 32) return f;
  const uint2 raw = *(const uint2 *)(c.feat16 + (size_t)t * 8 + (lane < 16 ? 0 : 4));
  f.x = __builtin_bit_cast(_Float16, (uint16_t)(raw.x & 0xffff));
  f.y = __builtin_bit_cast(_Float16, (uint16_t)(raw.x >> 16));
  if (lane < 16) {
    f.z = __builtin_bit_cast(_Float16, (uint16_t)(raw.y & 0xffff));
    f.w = __builtin_bit_cast(_Float16, (uint16_t)(raw.y >> 16));
  } else {
    f.z = (_Float16)feature_local_norm(c.seg2[tcol], c.seg_scale);
    f.w = (_Float16)c.progress;
  }
  return f;
}

}  // namespace mf
