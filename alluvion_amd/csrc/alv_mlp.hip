// alv_mlp.hip -- training and standalone inference of the learned
// warm-start network (see alv_mlp.hpp for the operand layout), and its C ABI.
//
// Training step (all on the device, one stream):
//   k_mlp_gather      random mini-batch, normalised inputs and targets
//   k_mlp_fwd_bwd     forward + backward through the activations, per tile
//                     of 16 samples, register-chained MFMAs; writes the bf16
//                     activations and deltas feature-major ([unit][sample])
//   k_mlp_wgrad       weight and bias gradients, MFMA with the batch as the
//                     contraction dimension, one partial per sample chunk
//   k_mlp_reduce      partials -> gradient, fixed order (deterministic)
//   k_mlp_adam        Adam on the FP32 master parameters
//   k_mlp_pack        FP32 parameters -> permuted bf16 operands
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "alv/alv.h"
#include "alv_mlp.hpp"

namespace alv {
int set_error(int code, const std::string& msg);  // alv_capi.hip
}

using namespace alv;

#define MLP_HIP(expr)                                                          \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return set_error(ALV_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace {

constexpr unsigned kChunk = 2048;  // samples per weight-gradient partial
constexpr unsigned kWgradJobs = 53;

__global__ __launch_bounds__(256) void k_mlp_pack(const float* __restrict__ P,
                                                  bf16x8* __restrict__ ops) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= kOpsTotal) return;
  const int lane = t & 63, g = lane >> 4, c = lane & 15;
  const int tile = t >> 6;  // index of the (m, s) pair inside the whole array
  bf16x8 v;
  float f[8];
  if (t < kOpsF2) {  // F1 [4][1]: W1[row][k], k = 8g + j (< 8 only for g = 0)
    const int m = tile, row = m * 16 + c;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = 8 * g + j;
      f[j] = k < kMlpIn ? P[kOffW1 + row * kMlpIn + k] : 0.0f;
    }
  } else if (t < kOpsF4) {  // F2, F3 [4][2]
    const bool l2 = t < kOpsF3;
    const int tt = (t - (l2 ? kOpsF2 : kOpsF3)) >> 6, m = tt >> 1, s = tt & 1;
    const int row = m * 16 + c, off = l2 ? kOffW2 : kOffW3;
#pragma unroll
    for (int j = 0; j < 8; ++j) f[j] = P[off + row * kMlpHid + hperm(s, g, j)];
  } else if (t < kOpsT4) {  // F4 [1][2]
    const int s = (t - kOpsF4) >> 6, row = c;
#pragma unroll
    for (int j = 0; j < 8; ++j)
      f[j] = row < kMlpOut ? P[kOffW4 + row * kMlpHid + hperm(s, g, j)] : 0.0f;
  } else if (t < kOpsT3) {  // T4 [4][1]: W4^T[in][o], o = hperm(0, g, j) < 2
    const int m = (t - kOpsT4) >> 6, row = m * 16 + c;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int o = hperm(0, g, j);
      f[j] = o < kMlpOut ? P[kOffW4 + o * kMlpHid + row] : 0.0f;
    }
  } else {  // T3, T2 [4][2]: W^T[in][o], o = hperm(s, g, j)
    const bool l3 = t < kOpsT2;
    const int tt = (t - (l3 ? kOpsT3 : kOpsT2)) >> 6, m = tt >> 1, s = tt & 1;
    const int row = m * 16 + c, off = l3 ? kOffW3 : kOffW2;
#pragma unroll
    for (int j = 0; j < 8; ++j) f[j] = P[off + hperm(s, g, j) * kMlpHid + row];
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) v[j] = (__bf16)f[j];
  ops[t] = v;
}

// SplitMix64 finaliser: sample index of (seed, step, i).
__device__ inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_mlp_gather(const float* __restrict__ X,
                                                    const float* __restrict__ Y, uint64_t n_data,
                                                    const float* __restrict__ norm, uint64_t seed,
                                                    unsigned step, unsigned B,
                                                    float* __restrict__ xn,
                                                    float* __restrict__ tn) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B) return;
  const uint64_t idx = mix64(seed ^ mix64((uint64_t(step) << 32) | i)) % n_data;
#pragma unroll
  for (int k = 0; k < kMlpIn; ++k)
    xn[size_t(i) * kMlpIn + k] = (X[idx * kMlpIn + k] - norm[k]) * norm[kMlpIn + k];
#pragma unroll
  for (int o = 0; o < kMlpOut; ++o)
    tn[size_t(i) * kMlpOut + o] =
        (Y[idx * kMlpOut + o] - norm[2 * kMlpIn + o]) / norm[2 * kMlpIn + 2 + o];
}

struct TrainBufs {
  __bf16* U;    // [8][B]
  __bf16* A1;   // [64][B]
  __bf16* A2;
  __bf16* A3;
  __bf16* DY;   // [2][B]
  __bf16* DZ3;  // [64][B]
  __bf16* DZ2;
  __bf16* DZ1;
};

__device__ inline void store_operands(__bf16* dst, const bf16x8 (&op)[2], int g, unsigned p,
                                      unsigned B) {
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int j = 0; j < 8; ++j) dst[size_t(hperm(s, g, j)) * B + p] = op[s][j];
}
__device__ inline void store_tiles(__bf16* dst, const f32x4 (&t)[4], int g, unsigned p,
                                   unsigned B) {
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int r = 0; r < 4; ++r) dst[size_t(m * 16 + 4 * g + r) * B + p] = (__bf16)t[m][r];
}
// delta tiles: round to bf16 once, so that the stored value and the operand
// of the next product are the same number
__device__ inline void round_tiles(f32x4 (&t)[4]) {
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int r = 0; r < 4; ++r) t[m][r] = float((__bf16)t[m][r]);
}

// Forward + backward of tiles of 16 samples (B a multiple of 16).  Every
// lane of every wave reaches the MFMAs; there is no early return.
__global__ __launch_bounds__(256) void k_mlp_fwd_bwd(MlpDev M, const float* __restrict__ xn,
                                                     const float* __restrict__ tn, unsigned B,
                                                     TrainBufs T, double* __restrict__ loss_part) {
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  const unsigned wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const unsigned n_waves = (gridDim.x * blockDim.x) >> 6;
  const unsigned n_tiles = B / 16;
  const bf16x8* ops = M.ops;
  const float scale = 2.0f / float(B);
  float loss = 0.0f;
  for (unsigned tile = wave; tile < n_tiles; tile += n_waves) {
    const unsigned p = tile * 16 + c;
    bf16x8 u;
#pragma unroll
    for (int k = 0; k < kMlpIn; ++k) u[k] = (__bf16)(g == 0 ? xn[size_t(p) * kMlpIn + k] : 0.0f);
    MlpTileState S;
    mlp_forward_tile(M, u, lane, S);
    // output delta (rows 0, 1 live in lanes of group 0)
    bf16x8 dyop;
#pragma unroll
    for (int j = 0; j < 8; ++j) dyop[j] = (__bf16)0.0f;
    if (g == 0) {
#pragma unroll
      for (int o = 0; o < kMlpOut; ++o) {
        const float e = S.y[o] - tn[size_t(p) * kMlpOut + o];
        loss += e * e;
        const __bf16 d = (__bf16)(scale * e);
        dyop[o] = d;
        T.DY[size_t(o) * B + p] = d;
      }
#pragma unroll
      for (int k = 0; k < kMlpIn; ++k) T.U[size_t(k) * B + p] = u[k];
    }
    store_operands(T.A1, S.a1, g, p, B);
    store_operands(T.A2, S.a2, g, p, B);
    store_operands(T.A3, S.a3, g, p, B);
    // layer 4 -> 3
    f32x4 d[4];
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      d[m] = mfma16x16x32(ops[kOpsT4 + m * 64 + lane], dyop, zero);
#pragma unroll
      for (int r = 0; r < 4; ++r) d[m][r] *= mlp_silu_grad(S.z3[m][r]);
    }
    round_tiles(d);
    store_tiles(T.DZ3, d, g, p, B);
    bf16x8 dop[2];
    tiles_to_operands(d, dop);
    // layer 3 -> 2
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      f32x4 acc = mfma16x16x32(ops[kOpsT3 + (m * 2 + 0) * 64 + lane], dop[0], zero);
      acc = mfma16x16x32(ops[kOpsT3 + (m * 2 + 1) * 64 + lane], dop[1], acc);
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r] *= mlp_silu_grad(S.z2[m][r]);
      d[m] = acc;
    }
    round_tiles(d);
    store_tiles(T.DZ2, d, g, p, B);
    tiles_to_operands(d, dop);
    // layer 2 -> 1
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      f32x4 acc = mfma16x16x32(ops[kOpsT2 + (m * 2 + 0) * 64 + lane], dop[0], zero);
      acc = mfma16x16x32(ops[kOpsT2 + (m * 2 + 1) * 64 + lane], dop[1], acc);
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r] *= mlp_silu_grad(S.z1[m][r]);
      d[m] = acc;
    }
    round_tiles(d);
    store_tiles(T.DZ1, d, g, p, B);
  }
  // block sum of the squared errors
  __shared__ float red[256];
  red[threadIdx.x] = loss;
  __syncthreads();
  for (unsigned w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// One wave per (job, chunk).  Jobs 0-39: weight tiles; 40-52: bias tiles.
struct WgradJob {
  int layer;   // 1..4
  int mt, nt;  // output-unit tile, input-unit tile (nt < 0: bias)
};
__device__ inline WgradJob wgrad_job(int j) {
  if (j < 4) return {1, j, 0};
  if (j < 20) return {2, (j - 4) / 4, (j - 4) % 4};
  if (j < 36) return {3, (j - 20) / 4, (j - 20) % 4};
  if (j < 40) return {4, 0, j - 36};
  if (j < 44) return {1, j - 40, -1};
  if (j < 48) return {2, j - 44, -1};
  if (j < 52) return {3, j - 48, -1};
  return {4, 0, -1};
}

__global__ __launch_bounds__(64) void k_mlp_wgrad(TrainBufs T, unsigned B,
                                                  float* __restrict__ partials) {
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  const WgradJob J = wgrad_job(blockIdx.x);
  const unsigned chunk = blockIdx.y, p_begin = chunk * kChunk;
  const __bf16* dz;
  const __bf16* ain;
  int n_out, n_in, w_off, b_off;
  switch (J.layer) {
    case 1: dz = T.DZ1; ain = T.U; n_out = kMlpHid; n_in = kMlpIn; w_off = kOffW1; b_off = kOffB1; break;
    case 2: dz = T.DZ2; ain = T.A1; n_out = kMlpHid; n_in = kMlpHid; w_off = kOffW2; b_off = kOffB2; break;
    case 3: dz = T.DZ3; ain = T.A2; n_out = kMlpHid; n_in = kMlpHid; w_off = kOffW3; b_off = kOffB3; break;
    default: dz = T.DY; ain = T.A3; n_out = kMlpOut; n_in = kMlpHid; w_off = kOffW4; b_off = kOffB4; break;
  }
  const int arow = J.mt * 16 + c;                       // output unit of this lane's A row
  const int brow = J.nt < 0 ? 0 : J.nt * 16 + c;        // input unit of this lane's B column
  const bool a_ok = arow < n_out, b_ok = J.nt < 0 || brow < n_in;
  bf16x8 ones, zeros;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    ones[j] = (__bf16)1.0f;
    zeros[j] = (__bf16)0.0f;
  }
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (unsigned p0 = p_begin; p0 < p_begin + kChunk; p0 += 32) {
    const unsigned p = p0 + 8 * g;
    const bf16x8 a = a_ok ? *reinterpret_cast<const bf16x8*>(dz + size_t(arow) * B + p) : zeros;
    const bf16x8 b = J.nt < 0 ? ones
                     : (b_ok ? *reinterpret_cast<const bf16x8*>(ain + size_t(brow) * B + p)
                             : zeros);
    acc = mfma16x16x32(a, b, acc);
  }
  float* out = partials + size_t(chunk) * kMlpParams;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = J.mt * 16 + 4 * g + r;
    if (row >= n_out) continue;
    if (J.nt < 0) {
      if (c == 0) out[b_off + row] = acc[r];
    } else {
      const int col = J.nt * 16 + c;
      if (col < n_in) out[w_off + row * n_in + col] = acc[r];
    }
  }
}

__global__ __launch_bounds__(256) void k_mlp_reduce(const float* __restrict__ partials,
                                                    unsigned n_chunks, float* __restrict__ grad) {
  const unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= unsigned(kMlpParams)) return;
  double s = 0.0;
  for (unsigned ch = 0; ch < n_chunks; ++ch) s += partials[size_t(ch) * kMlpParams + k];
  grad[k] = float(s);
}

__global__ __launch_bounds__(256) void k_mlp_adam(float* __restrict__ P, float* __restrict__ m1,
                                                  float* __restrict__ m2,
                                                  const float* __restrict__ grad, float lr,
                                                  float c1, float c2) {
  const unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= unsigned(kMlpParams)) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float gk = grad[k];
  const float a = b1 * m1[k] + (1.0f - b1) * gk;
  const float b = b2 * m2[k] + (1.0f - b2) * gk * gk;
  m1[k] = a;
  m2[k] = b;
  P[k] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// Standalone inference: raw features [n][8] -> de-normalised outputs [n][2].
__global__ __launch_bounds__(256) void k_mlp_forward(MlpDev M, const float* __restrict__ x,
                                                     size_t n, float* __restrict__ y) {
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  const size_t wave = (size_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const size_t n_waves = (size_t(gridDim.x) * blockDim.x) >> 6;
  const size_t n_tiles = (n + 15) / 16;
  for (size_t tile = wave; tile < n_tiles; tile += n_waves) {
    const size_t p = tile * 16 + c;
    const bool ok = g == 0 && p < n;
    bf16x8 u;
#pragma unroll
    for (int k = 0; k < kMlpIn; ++k)
      u[k] = (__bf16)(ok ? (x[p * kMlpIn + k] - M.norm[k]) * M.norm[kMlpIn + k] : 0.0f);
    const f32x4 o = mlp_infer_tile(M, u, lane);
    if (ok) {
      y[p * 2 + 0] = o[0] * M.norm[2 * kMlpIn + 2] + M.norm[2 * kMlpIn + 0];
      y[p * 2 + 1] = o[1] * M.norm[2 * kMlpIn + 3] + M.norm[2 * kMlpIn + 1];
    }
  }
}

}  // namespace

struct alv_mlp {
  hipStream_t stream = nullptr;
  float* params = nullptr;  // FP32 master
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  float* grad = nullptr;
  float* norm = nullptr;  // 20 floats
  bf16x8* ops = nullptr;
  unsigned adam_t = 0;
  MlpDev dev() const { return MlpDev{ops, params, norm}; }
};

namespace {
int pack(alv_mlp* m) {
  hipLaunchKernelGGL(k_mlp_pack, dim3((kOpsTotal + 255) / 256), dim3(256), 0, m->stream,
                     m->params, m->ops);
  MLP_HIP(hipGetLastError());
  return ALV_OK;
}

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

int alloc_train(alv_mlp* m, unsigned B, DevBuf& blob, TrainBufs& T, float** xn, float** tn,
                float** partials, double** loss_part, unsigned* fb_blocks) {
  const size_t per = size_t(B) * sizeof(__bf16);
  const unsigned n_chunks = B / kChunk;
  *fb_blocks = std::min(1024u, (B / 16 + 3) / 4);
  const size_t bytes = per * (8 + 64 * 3 + 2 + 64 * 3) + size_t(B) * (8 + 2) * sizeof(float) +
                       size_t(n_chunks) * kMlpParams * sizeof(float) +
                       size_t(*fb_blocks) * sizeof(double) + 16 * 256;
  MLP_HIP(hipMalloc(&blob.p, bytes));
  char* q = static_cast<char*>(blob.p);
  auto take = [&](size_t n) {
    char* r = q;
    q += (n + 255) / 256 * 256;
    return r;
  };
  T.U = reinterpret_cast<__bf16*>(take(per * 8));
  T.A1 = reinterpret_cast<__bf16*>(take(per * 64));
  T.A2 = reinterpret_cast<__bf16*>(take(per * 64));
  T.A3 = reinterpret_cast<__bf16*>(take(per * 64));
  T.DY = reinterpret_cast<__bf16*>(take(per * 2));
  T.DZ3 = reinterpret_cast<__bf16*>(take(per * 64));
  T.DZ2 = reinterpret_cast<__bf16*>(take(per * 64));
  T.DZ1 = reinterpret_cast<__bf16*>(take(per * 64));
  *xn = reinterpret_cast<float*>(take(size_t(B) * 8 * sizeof(float)));
  *tn = reinterpret_cast<float*>(take(size_t(B) * 2 * sizeof(float)));
  *partials = reinterpret_cast<float*>(take(size_t(n_chunks) * kMlpParams * sizeof(float)));
  *loss_part = reinterpret_cast<double*>(take(size_t(*fb_blocks) * sizeof(double)));
  return ALV_OK;
}

// gradient of the normalised batch (xn, tn already on the device)
int grad_batch(alv_mlp* m, const float* xn, const float* tn, unsigned B, TrainBufs& T,
               float* partials, double* loss_part, unsigned fb_blocks) {
  hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(fb_blocks), dim3(256), 0, m->stream, m->dev(), xn, tn,
                     B, T, loss_part);
  hipLaunchKernelGGL(k_mlp_wgrad, dim3(kWgradJobs, B / kChunk), dim3(64), 0, m->stream, T, B,
                     partials);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kMlpParams + 255) / 256), dim3(256), 0, m->stream,
                     partials, B / kChunk, m->grad);
  MLP_HIP(hipGetLastError());
  return ALV_OK;
}

int check_batch(unsigned B) {
  if (B == 0 || B % kChunk != 0)
    return set_error(ALV_ERR_ARG, "batch size must be a positive multiple of 2048");
  return ALV_OK;
}
}  // namespace

extern "C" {

int alv_mlp_num_params(void) { return kMlpParams; }

int alv_mlp_create(alv_mlp** out, unsigned long long seed) {
  if (!out) return set_error(ALV_ERR_ARG, "null output");
  alv_mlp* m = new alv_mlp();
  *out = nullptr;
  auto fail = [&](int rc) {
    alv_mlp_destroy(m);
    return rc;
  };
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&m->params, kMlpParams * sizeof(float)) != hipSuccess ||
      hipMalloc(&m->adam_m, kMlpParams * sizeof(float)) != hipSuccess ||
      hipMalloc(&m->adam_v, kMlpParams * sizeof(float)) != hipSuccess ||
      hipMalloc(&m->grad, kMlpParams * sizeof(float)) != hipSuccess ||
      hipMalloc(&m->norm, 20 * sizeof(float)) != hipSuccess ||
      hipMalloc(&m->ops, kOpsTotal * sizeof(bf16x8)) != hipSuccess)
    return fail(set_error(ALV_ERR_HIP, "alv_mlp_create: device allocation failed"));
  // uniform(-sqrt(6/(fan_in+fan_out)), +) weights, zero biases
  std::vector<float> P(kMlpParams, 0.0f);
  uint64_t s = seed;
  auto fill = [&](int off, int n_out, int n_in) {
    const double lim = std::sqrt(6.0 / double(n_in + n_out));
    for (int k = 0; k < n_out * n_in; ++k) {
      const double u = double(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
      P[off + k] = float((2.0 * u - 1.0) * lim);
    }
  };
  fill(kOffW1, kMlpHid, kMlpIn);
  fill(kOffW2, kMlpHid, kMlpHid);
  fill(kOffW3, kMlpHid, kMlpHid);
  fill(kOffW4, kMlpOut, kMlpHid);
  const float norm[20] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 1, 1};
  if (hipMemcpy(m->params, P.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(m->norm, norm, sizeof(norm), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(m->adam_m, 0, kMlpParams * sizeof(float)) != hipSuccess ||
      hipMemset(m->adam_v, 0, kMlpParams * sizeof(float)) != hipSuccess)
    return fail(set_error(ALV_ERR_HIP, "alv_mlp_create: upload failed"));
  int rc = pack(m);
  if (rc != ALV_OK) return fail(rc);
  if (hipStreamSynchronize(m->stream) != hipSuccess)
    return fail(set_error(ALV_ERR_HIP, "alv_mlp_create: pack failed"));
  *out = m;
  return ALV_OK;
}

int alv_mlp_destroy(alv_mlp* m) {
  if (!m) return ALV_OK;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (void* p : {static_cast<void*>(m->params), static_cast<void*>(m->adam_m),
                  static_cast<void*>(m->adam_v), static_cast<void*>(m->grad),
                  static_cast<void*>(m->norm), static_cast<void*>(m->ops)})
    if (p) (void)hipFree(p);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return ALV_OK;
}

int alv_mlp_set_params(alv_mlp* m, const float* host_params) {
  if (!m || !host_params) return set_error(ALV_ERR_ARG, "null argument");
  MLP_HIP(hipMemcpyAsync(m->params, host_params, kMlpParams * sizeof(float),
                         hipMemcpyHostToDevice, m->stream));
  int rc = pack(m);
  if (rc != ALV_OK) return rc;
  MLP_HIP(hipStreamSynchronize(m->stream));
  return ALV_OK;
}

int alv_mlp_get_params(alv_mlp* m, float* host_params) {
  if (!m || !host_params) return set_error(ALV_ERR_ARG, "null argument");
  MLP_HIP(hipMemcpyAsync(host_params, m->params, kMlpParams * sizeof(float),
                         hipMemcpyDeviceToHost, m->stream));
  MLP_HIP(hipStreamSynchronize(m->stream));
  return ALV_OK;
}

int alv_mlp_set_normalization(alv_mlp* m, const float* norm20) {
  if (!m || !norm20) return set_error(ALV_ERR_ARG, "null argument");
  for (int k = 0; k < 20; ++k)
    if (!std::isfinite(norm20[k])) return set_error(ALV_ERR_ARG, "non-finite normalisation");
  for (int o = 0; o < 2; ++o)
    if (norm20[18 + o] == 0.0f) return set_error(ALV_ERR_ARG, "zero output scale");
  MLP_HIP(hipMemcpyAsync(m->norm, norm20, 20 * sizeof(float), hipMemcpyHostToDevice, m->stream));
  MLP_HIP(hipStreamSynchronize(m->stream));
  return ALV_OK;
}

int alv_mlp_get_normalization(alv_mlp* m, float* norm20) {
  if (!m || !norm20) return set_error(ALV_ERR_ARG, "null argument");
  MLP_HIP(hipMemcpyAsync(norm20, m->norm, 20 * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MLP_HIP(hipStreamSynchronize(m->stream));
  return ALV_OK;
}

int alv_mlp_fit_normalization(alv_mlp* m, const float* dev_x, const float* dev_y, size_t n) {
  if (!m || !dev_x || !dev_y || n == 0) return set_error(ALV_ERR_ARG, "bad argument");
  // mean and standard deviation of every input and target over a strided
  // sample of at most 2^20 pairs
  const size_t stride = std::max<size_t>(1, n >> 20);
  const size_t ns = (n + stride - 1) / stride;
  std::vector<float> hx(n * kMlpIn), hy(n * kMlpOut);
  MLP_HIP(hipMemcpy(hx.data(), dev_x, n * kMlpIn * sizeof(float), hipMemcpyDeviceToHost));
  MLP_HIP(hipMemcpy(hy.data(), dev_y, n * kMlpOut * sizeof(float), hipMemcpyDeviceToHost));
  float norm[20];
  auto stats = [&](const std::vector<float>& v, int dim, int k, double& mean, double& sd) {
    double s = 0, s2 = 0;
    for (size_t i = 0; i < n; i += stride) s += v[i * dim + k];
    mean = s / double(ns);
    for (size_t i = 0; i < n; i += stride) {
      const double d = v[i * dim + k] - mean;
      s2 += d * d;
    }
    sd = std::sqrt(s2 / double(ns));
  };
  for (int k = 0; k < kMlpIn; ++k) {
    double mu, sd;
    stats(hx, kMlpIn, k, mu, sd);
    norm[k] = float(mu);
    norm[kMlpIn + k] = sd > 0 && std::isfinite(sd) ? float(1.0 / sd) : 1.0f;
  }
  for (int o = 0; o < kMlpOut; ++o) {
    double mu, sd;
    stats(hy, kMlpOut, o, mu, sd);
    norm[16 + o] = float(mu);
    norm[18 + o] = sd > 0 && std::isfinite(sd) ? float(sd) : 1.0f;
  }
  return alv_mlp_set_normalization(m, norm);
}

int alv_mlp_forward(alv_mlp* m, const float* dev_x, size_t n, float* dev_y) {
  if (!m || (n && (!dev_x || !dev_y))) return set_error(ALV_ERR_ARG, "bad argument");
  if (!n) return ALV_OK;
  const size_t tiles = (n + 15) / 16;
  const unsigned blocks = unsigned(std::min<size_t>(4096, (tiles + 3) / 4));
  hipLaunchKernelGGL(k_mlp_forward, dim3(blocks), dim3(256), 0, m->stream, m->dev(), dev_x, n,
                     dev_y);
  MLP_HIP(hipGetLastError());
  MLP_HIP(hipStreamSynchronize(m->stream));
  return ALV_OK;
}

int alv_mlp_loss_grad(alv_mlp* m, const float* dev_xn, const float* dev_tn, size_t n,
                      double* loss, float* host_grad) {
  if (!m || !dev_xn || !dev_tn || !loss || !host_grad) return set_error(ALV_ERR_ARG, "null argument");
  int rc = check_batch(unsigned(n));
  if (rc != ALV_OK) return rc;
  const unsigned B = unsigned(n);
  DevBuf blob;
  TrainBufs T;
  float *xn, *tn, *partials;
  double* loss_part;
  unsigned fb;
  rc = alloc_train(m, B, blob, T, &xn, &tn, &partials, &loss_part, &fb);
  if (rc != ALV_OK) return rc;
  rc = grad_batch(m, dev_xn, dev_tn, B, T, partials, loss_part, fb);
  if (rc != ALV_OK) return rc;
  std::vector<double> lp(fb);
  MLP_HIP(hipMemcpyAsync(host_grad, m->grad, kMlpParams * sizeof(float), hipMemcpyDeviceToHost,
                         m->stream));
  MLP_HIP(hipMemcpyAsync(lp.data(), loss_part, fb * sizeof(double), hipMemcpyDeviceToHost,
                         m->stream));
  MLP_HIP(hipStreamSynchronize(m->stream));
  double L = 0;
  for (double v : lp) L += v;
  *loss = L / double(B);
  return ALV_OK;
}

int alv_mlp_train(alv_mlp* m, const float* dev_x, const float* dev_y, size_t n,
                  unsigned batch, unsigned steps, float lr, unsigned long long seed,
                  float* host_loss) {
  if (!m || !dev_x || !dev_y || n == 0) return set_error(ALV_ERR_ARG, "bad argument");
  int rc = check_batch(batch);
  if (rc != ALV_OK) return rc;
  DevBuf blob;
  TrainBufs T;
  float *xn, *tn, *partials;
  double* loss_part;
  unsigned fb;
  rc = alloc_train(m, batch, blob, T, &xn, &tn, &partials, &loss_part, &fb);
  if (rc != ALV_OK) return rc;
  std::vector<double> lp(fb);
  for (unsigned st = 0; st < steps; ++st) {
    hipLaunchKernelGGL(k_mlp_gather, dim3((batch + 255) / 256), dim3(256), 0, m->stream, dev_x,
                       dev_y, uint64_t(n), m->norm, uint64_t(seed), st, batch, xn, tn);
    rc = grad_batch(m, xn, tn, batch, T, partials, loss_part, fb);
    if (rc != ALV_OK) return rc;
    ++m->adam_t;
    const float c1 = 1.0f - std::pow(0.9f, float(m->adam_t));
    const float c2 = 1.0f - std::pow(0.999f, float(m->adam_t));
    hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, m->stream,
                       m->params, m->adam_m, m->adam_v, m->grad, lr, c1, c2);
    rc = pack(m);
    if (rc != ALV_OK) return rc;
    if (host_loss) {
      MLP_HIP(hipMemcpyAsync(lp.data(), loss_part, fb * sizeof(double), hipMemcpyDeviceToHost,
                             m->stream));
      MLP_HIP(hipStreamSynchronize(m->stream));
      double L = 0;
      for (double v : lp) L += v;
      host_loss[st] = float(L / double(batch));
    }
  }
  MLP_HIP(hipStreamSynchronize(m->stream));
  return ALV_OK;
}

}  // extern "C"

namespace alv {
MlpDev mlp_device_view(const alv_mlp* m) { return m ? m->dev() : MlpDev{nullptr, nullptr, nullptr}; }
}
