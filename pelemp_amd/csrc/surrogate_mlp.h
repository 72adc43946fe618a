// surrogate_mlp.h -- bfloat16 MFMA inference of the surrogate network on
// gfx950, one wavefront evaluating the 64 droplets of its lanes.
//
// Formulation.  Each linear layer is computed transposed, D = W * H: the
// weights are the A operand (rows = output features), the activations the B
// operand (columns = samples), with v_mfma_f32_32x32x16_bf16 (A 32x16,
// B 16x32, D 32x32 in FP32).  Operand layout of that instruction (lane l,
// element v of the 8-element bf16 vector, register r of the 16-float
// accumulator):
//   A[i][k]: lane i + 32*(k/8), element k%8
//   B[k][j]: lane j + 32*(k/8), element k%8
//   D[i][j]: lane j + 32*((i/4)%2), register 4*(i/8) + i%4
// The wave's 64 droplets form two blocks of 32 columns.  Layer inputs:
// lanes 0-31 hold the 8 inputs of their droplet in k = 0..7, lanes 32-63 in
// k = 8..15, so one B register set serves both blocks; block 0 uses the
// weights in k = 0..7 (zeros in 8..15), block 1 the reverse.  Between
// hidden layers the FP32 accumulator of each lane holds, for its sample,
// the features i(r) above; packing registers 8s..8s+7 of the two output
// tiles as the B operand of K-step s feeds feature
//   F(s, h, v) = 16 s + 8 (v/4) + 4 h + v%4      (h = lane/32)
// into k-slot 8h + v, so the next layer's weights are stored with their
// columns permuted accordingly and no data moves between lanes.  The output
// layer (3 rows, padded to 32) puts block 0's results in rows 0-2 and block
// 1's in rows 4-6, so each lane receives the outputs of its own droplet.
// Biases enter as the accumulator's initial value.
//
// All weight fragments live in LDS (36 KiB, ds_read_b128 per MFMA, shared
// by the two blocks); accumulation is FP32, activations are rounded to
// bfloat16 (round to nearest even) at each layer input.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "surrogate.h"

namespace pelemp {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

constexpr int kSurrFrag0 = 0;    // input layer: (m, blk) -> m*2 + blk
constexpr int kSurrFragH = 4;    // hidden layer L=1..3: 4 + (L-1)*8 + m*4 + s
constexpr int kSurrFragO = 28;   // output layer: 28 + blk*4 + s
constexpr int kSurrFrags = 36;

// LDS image of the packed network (also its device-memory layout).
struct SurrWeights {
  uint4 frag[kSurrFrags * 64];   // 64 lanes x 8 bf16 per fragment
  float bias[4 * kSurrHid];      // hidden-layer biases b_0..b_3
  float bout[4];                 // output bias b_4 (3 used)
};
constexpr int kSurrWeightWords = static_cast<int>(sizeof(SurrWeights) / 4);

// ---- host-side packing ----------------------------------------------------
inline uint16_t bf16_rne(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  const uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(r >> 16);
}

// feature index held in k-slot (h, v) of K-step s (see header comment)
inline int surr_kfeature(int s, int h, int v) { return 16 * s + 8 * (v / 4) + 4 * h + (v % 4); }

inline void surr_pack(const float* P, SurrWeights& W) {
  auto put = [&](int f, int lane, int v, float val) {
    uint16_t* base = reinterpret_cast<uint16_t*>(&W.frag[f * 64 + lane]);
    base[v] = bf16_rne(val);
  };
  for (int f = 0; f < kSurrFrags; ++f)
    for (int lane = 0; lane < 64; ++lane) W.frag[f * 64 + lane] = make_uint4(0, 0, 0, 0);
  // input layer
  for (int m = 0; m < 2; ++m)
    for (int blk = 0; blk < 2; ++blk)
      for (int lane = 0; lane < 64; ++lane) {
        const int h = lane / 32, row = 32 * m + lane % 32;
        if (h != blk) continue;
        for (int v = 0; v < 8; ++v) put(kSurrFrag0 + m * 2 + blk, lane, v, P[surr_w_off(0) + row * kSurrIn + v]);
      }
  // hidden layers 1..3
  for (int L = 1; L <= 3; ++L)
    for (int m = 0; m < 2; ++m)
      for (int s = 0; s < 4; ++s)
        for (int lane = 0; lane < 64; ++lane) {
          const int h = lane / 32, row = 32 * m + lane % 32;
          for (int v = 0; v < 8; ++v)
            put(kSurrFragH + (L - 1) * 8 + m * 4 + s, lane, v,
                P[surr_w_off(L) + row * kSurrHid + surr_kfeature(s, h, v)]);
        }
  // output layer
  for (int blk = 0; blk < 2; ++blk)
    for (int s = 0; s < 4; ++s)
      for (int lane = 0; lane < 64; ++lane) {
        const int h = lane / 32, i = lane % 32;
        const int orow = (blk == 0) ? i : i - 4;
        if (orow < 0 || orow >= kSurrOut) continue;
        for (int v = 0; v < 8; ++v)
          put(kSurrFragO + blk * 4 + s, lane, v, P[surr_w_off(4) + orow * kSurrHid + surr_kfeature(s, h, v)]);
      }
  for (int L = 0; L < 4; ++L)
    for (int o = 0; o < kSurrHid; ++o) W.bias[L * kSurrHid + o] = P[surr_b_off(L) + o];
  for (int q = 0; q < 4; ++q) W.bout[q] = (q < kSurrOut) ? P[surr_b_off(4) + q] : 0.f;
}

// ---- device side ------------------------------------------------------------
#if defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC__)

// Cooperative copy of the packed network into LDS (all threads of the block).
__device__ inline void surr_stage(SurrWeights& dst, const SurrWeights* __restrict__ src, int t, int nthreads) {
  const uint4* s = reinterpret_cast<const uint4*>(src);
  uint4* d = reinterpret_cast<uint4*>(&dst);
  constexpr int n16 = static_cast<int>(sizeof(SurrWeights) / 16);
  for (int i = t; i < n16; i += nthreads) d[i] = s[i];
}

__device__ __attribute__((always_inline)) inline bf16x8_t surr_frag(const SurrWeights& W, int f, int lane) {
  return __builtin_bit_cast(bf16x8_t, W.frag[f * 64 + lane]);
}

// accumulator initialised with the bias of tile m (rows 32m..32m+31)
__device__ __attribute__((always_inline)) inline f32x16_t surr_bias_tile(const float* b, int m, int h) {
  f32x16_t c;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const float4 v = *reinterpret_cast<const float4*>(&b[32 * m + 8 * q + 4 * h]);
    c[4 * q + 0] = v.x;
    c[4 * q + 1] = v.y;
    c[4 * q + 2] = v.z;
    c[4 * q + 3] = v.w;
  }
  return c;
}

// The hidden layers' activation, hard swish: z clamp(z/6 + 1/2, 0, 1) (a
// multiply-add, a median of three and a multiply; no transcendental).  The
// same function (one definition) in training and inference.
__device__ __attribute__((always_inline)) inline float surr_act(float z) {
  const float g = __fmaf_rn(z, 1.f / 6.f, 0.5f);
  return z * __builtin_amdgcn_fmed3f(g, 0.f, 1.f);
}

// Activation of the two output tiles, packed as the 4 K-step B operands
__device__ __attribute__((always_inline)) inline void surr_act_pack(const f32x16_t* acc, bf16x8_t* B) {
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int r = 0; r < 16; ++r) B[2 * m + r / 8][r % 8] = static_cast<__bf16>(surr_act(acc[m][r]));
}

// Evaluate the network for the droplet of every lane.  Must be called by
// all 64 lanes of the wavefront (MFMA is a wavefront instruction); x are the
// normalised inputs, y the normalised outputs.
__device__ inline void surr_mlp(const SurrWeights& W, const float* x, float* y) {
  const int lane = threadIdx.x & 63;
  const int h = lane >> 5;
  bf16x8_t bx;
#pragma unroll
  for (int v = 0; v < 8; ++v) bx[v] = static_cast<__bf16>(x[v]);
  bf16x8_t B[2][4];
  // input layer
#pragma unroll
  for (int blk = 0; blk < 2; ++blk) {
    f32x16_t acc[2];
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      acc[m] = surr_bias_tile(&W.bias[0], m, h);
      acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(surr_frag(W, kSurrFrag0 + m * 2 + blk, lane), bx, acc[m], 0, 0,
                                                       0);
    }
    surr_act_pack(acc, B[blk]);
  }
  // hidden layers (each weight fragment feeds both blocks)
#pragma unroll
  for (int L = 1; L <= 3; ++L) {
    f32x16_t acc[2][2];
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      acc[0][m] = surr_bias_tile(&W.bias[L * kSurrHid], m, h);
      acc[1][m] = acc[0][m];
    }
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const bf16x8_t a = surr_frag(W, kSurrFragH + (L - 1) * 8 + m * 4 + s, lane);
        acc[0][m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, B[0][s], acc[0][m], 0, 0, 0);
        acc[1][m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, B[1][s], acc[1][m], 0, 0, 0);
      }
    surr_act_pack(acc[0], B[0]);
    surr_act_pack(acc[1], B[1]);
  }
  // output layer
  f32x16_t o[2];
#pragma unroll
  for (int blk = 0; blk < 2; ++blk) {
#pragma unroll
    for (int r = 0; r < 16; ++r) o[blk][r] = (r < kSurrOut && h == blk) ? W.bout[r] : 0.f;
#pragma unroll
    for (int s = 0; s < 4; ++s)
      o[blk] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(surr_frag(W, kSurrFragO + blk * 4 + s, lane), B[blk][s], o[blk],
                                                       0, 0, 0);
  }
#pragma unroll
  for (int q = 0; q < kSurrOut; ++q) y[q] = h ? o[1][q] : o[0][q];
}

#endif

}  // namespace pelemp
