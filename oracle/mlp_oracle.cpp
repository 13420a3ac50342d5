// mlp_oracle.cpp -- CPU restatement of the learned warm-start network.
//
// TEST INFRASTRUCTURE ONLY (part of liboracle.so; see alv_oracle.h).
//
// Network: 8 inputs -> 64 -> 64 -> 64 (SiLU) -> 2 outputs, the layer sizes
// and activation of the warm-start model.  The reference has no such model;
// this file fixes its arithmetic so that the device kernels can be checked:
//
//   parameters (FP32, flat): W1[64][8] b1[64] W2[64][64] b2[64] W3[64][64]
//                            b3[64] W4[2][64] b4[2]                (9026)
//   u  = q(x)                       x: normalised inputs
//   z1 = q(W1) u  + b1,  a1 = q(silu(z1))
//   z2 = q(W2) a1 + b2,  a2 = q(silu(z2))
//   z3 = q(W3) a2 + b3,  a3 = q(silu(z3))
//   y  = q(W4) a3 + b4
//   L  = (1/n) sum_p sum_o (y_o - t_o)^2
//   dy = q(2 (y - t) / n)
//   dz3 = q((q(W4)^T dy) * silu'(z3)), dz2, dz1 likewise
//   dW_l = sum_p dz_l a_{l-1}^T,  db_l = sum_p dz_l
//
// q() rounds to bfloat16 (round to nearest even) when `quantize` is set and
// is the identity otherwise; everything else is FP64.  With quantize = 0 the
// gradient is the exact gradient of L, which a finite-difference test pins.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "alv_oracle.h"

namespace {

constexpr int kIn = 8, kHid = 64, kOut = 2;
constexpr int oW1 = 0, oB1 = oW1 + kHid * kIn, oW2 = oB1 + kHid, oB2 = oW2 + kHid * kHid,
              oW3 = oB2 + kHid, oB3 = oW3 + kHid * kHid, oW4 = oB3 + kHid,
              oB4 = oW4 + kOut * kHid, kParams = oB4 + kOut;
static_assert(kParams == 9026, "parameter count");

double bf16(double v) {
  float f = static_cast<float>(v);
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return f;
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  std::memcpy(&f, &u, 4);
  return f;
}

struct Net {
  int quantize;
  std::vector<double> w;  // parameters, weights already passed through q()
  double q(double v) const { return quantize ? bf16(v) : v; }
  explicit Net(const float* params, int quant) : quantize(quant), w(kParams) {
    for (int k = 0; k < kParams; ++k) w[k] = params[k];
    auto qw = [&](int off, int cnt) {
      for (int k = 0; k < cnt; ++k) w[off + k] = q(w[off + k]);
    };
    qw(oW1, kHid * kIn);
    qw(oW2, kHid * kHid);
    qw(oW3, kHid * kHid);
    qw(oW4, kOut * kHid);
  }
};

double silu(double z) { return z / (1.0 + std::exp(-z)); }
double silu_grad(double z) {
  double s = 1.0 / (1.0 + std::exp(-z));
  return s * (1.0 + z * (1.0 - s));
}

struct Sample {
  double u[kIn], z1[kHid], a1[kHid], z2[kHid], a2[kHid], z3[kHid], a3[kHid], y[kOut];
};

void forward(const Net& N, const float* x, Sample& S) {
  const double* w = N.w.data();
  for (int k = 0; k < kIn; ++k) S.u[k] = N.q(x[k]);
  for (int o = 0; o < kHid; ++o) {
    double acc = 0;
    for (int k = 0; k < kIn; ++k) acc += w[oW1 + o * kIn + k] * S.u[k];
    S.z1[o] = acc + w[oB1 + o];
    S.a1[o] = N.q(silu(S.z1[o]));
  }
  for (int o = 0; o < kHid; ++o) {
    double acc = 0;
    for (int k = 0; k < kHid; ++k) acc += w[oW2 + o * kHid + k] * S.a1[k];
    S.z2[o] = acc + w[oB2 + o];
    S.a2[o] = N.q(silu(S.z2[o]));
  }
  for (int o = 0; o < kHid; ++o) {
    double acc = 0;
    for (int k = 0; k < kHid; ++k) acc += w[oW3 + o * kHid + k] * S.a2[k];
    S.z3[o] = acc + w[oB3 + o];
    S.a3[o] = N.q(silu(S.z3[o]));
  }
  for (int o = 0; o < kOut; ++o) {
    double acc = 0;
    for (int k = 0; k < kHid; ++k) acc += w[oW4 + o * kHid + k] * S.a3[k];
    S.y[o] = acc + w[oB4 + o];
  }
}

}  // namespace

extern "C" {

int orc_mlp_num_params(void) { return kParams; }

int orc_mlp_forward(const float* params, size_t n, const float* x, int quantize, double* y) {
  if (!params || (n && (!x || !y))) return -1;
  Net N(params, quantize);
#pragma omp parallel for schedule(static)
  for (long p = 0; p < long(n); ++p) {
    Sample S;
    forward(N, x + size_t(p) * kIn, S);
    for (int o = 0; o < kOut; ++o) y[size_t(p) * kOut + o] = S.y[o];
  }
  return 0;
}

int orc_mlp_loss_grad(const float* params, size_t n, const float* x, const float* t,
                      int quantize, double* loss, double* grad) {
  if (!params || !grad || !loss || n == 0 || !x || !t) return -1;
  Net N(params, quantize);
  const double* w = N.w.data();
  std::vector<double> g(kParams, 0.0);
  double L = 0;
  // serial over samples: the sums are in sample order
  for (size_t p = 0; p < n; ++p) {
    Sample S;
    forward(N, x + p * kIn, S);
    double dy[kOut], dz3[kHid], dz2[kHid], dz1[kHid];
    for (int o = 0; o < kOut; ++o) {
      const double e = S.y[o] - t[p * kOut + o];
      L += e * e;
      dy[o] = N.q(2.0 * e / double(n));
    }
    for (int o = 0; o < kOut; ++o) {
      for (int k = 0; k < kHid; ++k) g[oW4 + o * kHid + k] += dy[o] * S.a3[k];
      g[oB4 + o] += dy[o];
    }
    for (int k = 0; k < kHid; ++k) {
      double acc = 0;
      for (int o = 0; o < kOut; ++o) acc += w[oW4 + o * kHid + k] * dy[o];
      dz3[k] = N.q(acc * silu_grad(S.z3[k]));
    }
    for (int o = 0; o < kHid; ++o) {
      for (int k = 0; k < kHid; ++k) g[oW3 + o * kHid + k] += dz3[o] * S.a2[k];
      g[oB3 + o] += dz3[o];
    }
    for (int k = 0; k < kHid; ++k) {
      double acc = 0;
      for (int o = 0; o < kHid; ++o) acc += w[oW3 + o * kHid + k] * dz3[o];
      dz2[k] = N.q(acc * silu_grad(S.z2[k]));
    }
    for (int o = 0; o < kHid; ++o) {
      for (int k = 0; k < kHid; ++k) g[oW2 + o * kHid + k] += dz2[o] * S.a1[k];
      g[oB2 + o] += dz2[o];
    }
    for (int k = 0; k < kHid; ++k) {
      double acc = 0;
      for (int o = 0; o < kHid; ++o) acc += w[oW2 + o * kHid + k] * dz2[o];
      dz1[k] = N.q(acc * silu_grad(S.z1[k]));
    }
    for (int o = 0; o < kHid; ++o) {
      for (int k = 0; k < kIn; ++k) g[oW1 + o * kIn + k] += dz1[o] * S.u[k];
      g[oB1 + o] += dz1[o];
    }
  }
  *loss = L / double(n);
  std::memcpy(grad, g.data(), kParams * sizeof(double));
  return 0;
}

}  // extern "C"
