// alv_mlp.hpp -- the learned warm-start network on CDNA4 matrix cores.
//
// Network: 8 inputs -> 64 -> 64 -> 64 (SiLU) -> 2 outputs.  Weights and
// activations are bfloat16, accumulation is FP32 (v_mfma_f32_16x16x32_bf16),
// biases are FP32 and enter as the accumulator's initial value.
//
// The layers are evaluated transposed, Z^T = W Z_prev^T: weights are the A
// operand (rows = output units), a tile of 16 samples is the B operand
// (columns = samples).  An MFMA result tile leaves lane l = 16 g + c holding
// rows 4g..4g+3 of column c.  Two such tiles (rows m = 2s and 2s+1) give that
// lane exactly the 8 values that the next layer's B operand needs for k-step
// s if the contraction index of that k-step is permuted as
//
//   hidden unit of (k-step s, lane group g, element j) =
//       (2s + j/4) * 16 + 4g + (j mod 4)                    (hperm)
//
// and the weights are stored with the same permutation.  So activations stay
// in registers from layer to layer with no LDS round trip or lane shuffle.
// The backward pass chains the transposed weights the same way.
#pragma once

#include <hip/hip_runtime.h>

struct alv_mlp;

namespace alv {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMlpIn = 8, kMlpHid = 64, kMlpOut = 2;
// flat FP32 parameter layout: W1[64][8] b1 W2[64][64] b2 W3[64][64] b3 W4[2][64] b4
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kMlpHid * kMlpIn, kOffW2 = kOffB1 + kMlpHid,
              kOffB2 = kOffW2 + kMlpHid * kMlpHid, kOffW3 = kOffB2 + kMlpHid,
              kOffB3 = kOffW3 + kMlpHid * kMlpHid, kOffW4 = kOffB3 + kMlpHid,
              kOffB4 = kOffW4 + kMlpOut * kMlpHid, kMlpParams = kOffB4 + kMlpOut;
static_assert(kMlpParams == 9026, "parameter count");

// Packed operand arrays (bf16x8 per lane, [tile][k-step][64 lanes]).
//   F1 [4][1]  F2 [4][2]  F3 [4][2]  F4 [1][2]    forward weights
//   T4 [4][1]  T3 [4][2]  T2 [4][2]               transposed weights (backward)
constexpr int kOpsF1 = 0, kOpsF2 = kOpsF1 + 4 * 64, kOpsF3 = kOpsF2 + 8 * 64,
              kOpsF4 = kOpsF3 + 8 * 64, kOpsT4 = kOpsF4 + 2 * 64, kOpsT3 = kOpsT4 + 4 * 64,
              kOpsT2 = kOpsT3 + 8 * 64, kOpsTotal = kOpsT2 + 8 * 64;

__host__ __device__ inline int hperm(int s, int g, int j) {
  return (2 * s + (j >> 2)) * 16 + 4 * g + (j & 3);
}

// Device view of one network (all pointers device memory).
struct MlpDev {
  const bf16x8* ops;     // kOpsTotal packed operands
  const float* params;   // FP32 master parameters (biases are read from here)
  const float* norm;     // [8] in_shift, [8] in_scale, [2] out_shift, [2] out_scale
};

__device__ inline f32x4 mfma16x16x32(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Accumulator initialised with the bias of rows m*16 + 4g + r (0 past nrows).
__device__ inline f32x4 bias_tile(const float* b, int m, int g, int nrows) {
  f32x4 c;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = m * 16 + 4 * g + r;
    c[r] = row < nrows ? b[row] : 0.0f;
  }
  return c;
}

__device__ inline float mlp_silu(float z) { return z / (1.0f + expf(-z)); }
// The same with the hardware exponential and reciprocal (about 1 ulp each;
// the network computes in bfloat16, rounding 2^-8), for the per-step
// inference; training keeps the correctly rounded form.
__device__ inline float mlp_silu_fast(float z) {
  return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}
__device__ inline float mlp_silu_grad(float z) {
  const float s = 1.0f / (1.0f + expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// Four result tiles (64 rows) -> the two B operands of the next layer.
__device__ inline void tiles_to_operands(const f32x4 (&t)[4], bf16x8 (&op)[2]) {
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      op[s][j] = (__bf16)t[2 * s][j];
      op[s][4 + j] = (__bf16)t[2 * s + 1][j];
    }
}

template <bool kFast = false>
__device__ inline void silu_tiles(const f32x4 (&z)[4], f32x4 (&a)[4]) {
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int r = 0; r < 4; ++r) a[m][r] = kFast ? mlp_silu_fast(z[m][r]) : mlp_silu(z[m][r]);
}

// One hidden layer: z = W x + b for 64 outputs from two k-steps of input.
__device__ inline void hidden_layer(const bf16x8* W, const float* b, const bf16x8 (&x)[2],
                                    int lane, int g, f32x4 (&z)[4]) {
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    f32x4 acc = bias_tile(b, m, g, kMlpHid);
    acc = mfma16x16x32(W[(m * 2 + 0) * 64 + lane], x[0], acc);
    acc = mfma16x16x32(W[(m * 2 + 1) * 64 + lane], x[1], acc);
    z[m] = acc;
  }
}

// Forward pass of one tile of 16 samples.  u: lanes 0-15 hold the 8
// normalised inputs of sample (lane mod 16) as bf16, other lanes zero.
// Result: lanes 0-15 hold y[0], y[1] of their sample in out[0], out[1].
// The pre-activations and the bf16 activation operands are returned for the
// backward pass.
struct MlpTileState {
  f32x4 z1[4], z2[4], z3[4];
  bf16x8 a1[2], a2[2], a3[2];
  f32x4 y;
};

__device__ inline void mlp_forward_tile(const MlpDev& M, bf16x8 u, int lane, MlpTileState& S) {
  const int g = lane >> 4;
  const bf16x8* ops = M.ops;
  const float* P = M.params;
#pragma unroll
  for (int m = 0; m < 4; ++m)
    S.z1[m] = mfma16x16x32(ops[kOpsF1 + m * 64 + lane], u, bias_tile(P + kOffB1, m, g, kMlpHid));
  f32x4 a[4];
  silu_tiles(S.z1, a);
  tiles_to_operands(a, S.a1);
  hidden_layer(ops + kOpsF2, P + kOffB2, S.a1, lane, g, S.z2);
  silu_tiles(S.z2, a);
  tiles_to_operands(a, S.a2);
  hidden_layer(ops + kOpsF3, P + kOffB3, S.a2, lane, g, S.z3);
  silu_tiles(S.z3, a);
  tiles_to_operands(a, S.a3);
  f32x4 y = bias_tile(P + kOffB4, 0, g, kMlpOut);
  y = mfma16x16x32(ops[kOpsF4 + 0 * 64 + lane], S.a3[0], y);
  y = mfma16x16x32(ops[kOpsF4 + 1 * 64 + lane], S.a3[1], y);
  S.y = y;
}

// Lightweight inference: same arithmetic, keeps only the outputs.
__device__ inline f32x4 mlp_infer_tile(const MlpDev& M, bf16x8 u, int lane) {
  MlpTileState S;
  mlp_forward_tile(M, u, lane, S);
  return S.y;
}

// Wave-level inference for 64 samples held one per lane (raw features f).
// Every lane of the wave must call it (MFMA is wave-wide).  Returns the
// de-normalised outputs of this lane's sample.  The four 16-sample tiles go
// through each layer together, so every weight operand is loaded once per
// wave; the arithmetic per tile is mlp_forward_tile's.
__device__ inline void mlp_infer_wave(const MlpDev& M, const float (&f)[kMlpIn], float& out0,
                                      float& out1) {
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  const bf16x8* ops = M.ops;
  const float* P = M.params;
  float xn[kMlpIn];
#pragma unroll
  for (int k = 0; k < kMlpIn; ++k) xn[k] = (f[k] - M.norm[k]) * M.norm[kMlpIn + k];
  bf16x8 u[4];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int k = 0; k < kMlpIn; ++k) {
      const float v = __shfl(xn[k], t * 16 + c, 64);
      u[t][k] = (__bf16)(g == 0 ? v : 0.0f);
    }
  f32x4 z[4][4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const bf16x8 w = ops[kOpsF1 + m * 64 + lane];
    const f32x4 b = bias_tile(P + kOffB1, m, g, kMlpHid);
#pragma unroll
    for (int t = 0; t < 4; ++t) z[t][m] = mfma16x16x32(w, u[t], b);
  }
  bf16x8 a[4][2];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    f32x4 act[4];
    silu_tiles<true>(z[t], act);
    tiles_to_operands(act, a[t]);
  }
  const int F[2] = {kOpsF2, kOpsF3}, B[2] = {kOffB2, kOffB3};
#pragma unroll
  for (int layer = 0; layer < 2; ++layer) {
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const bf16x8 w0 = ops[F[layer] + (m * 2 + 0) * 64 + lane];
      const bf16x8 w1 = ops[F[layer] + (m * 2 + 1) * 64 + lane];
      const f32x4 b = bias_tile(P + B[layer], m, g, kMlpHid);
#pragma unroll
      for (int t = 0; t < 4; ++t) z[t][m] = mfma16x16x32(w1, a[t][1], mfma16x16x32(w0, a[t][0], b));
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      f32x4 act[4];
      silu_tiles<true>(z[t], act);
      tiles_to_operands(act, a[t]);
    }
  }
  const bf16x8 w0 = ops[kOpsF4 + 0 * 64 + lane], w1 = ops[kOpsF4 + 1 * 64 + lane];
  const f32x4 b = bias_tile(P + kOffB4, 0, g, kMlpOut);
  float r0 = 0.0f, r1 = 0.0f;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const f32x4 y = mfma16x16x32(w1, a[t][1], mfma16x16x32(w0, a[t][0], b));
    // lanes 0-15 hold the outputs of samples t*16 + lane
    const float y0 = __shfl(y[0], c, 64), y1 = __shfl(y[1], c, 64);
    if (g == t) {
      r0 = y0;
      r1 = y1;
    }
  }
  out0 = r0 * M.norm[2 * kMlpIn + 2] + M.norm[2 * kMlpIn + 0];
  out1 = r1 * M.norm[2 * kMlpIn + 3] + M.norm[2 * kMlpIn + 1];
}

// device view of a network object (null network -> null pointers)
MlpDev mlp_device_view(const alv_mlp* m);

}  // namespace alv
