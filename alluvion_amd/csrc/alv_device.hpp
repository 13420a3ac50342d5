// alv_device.hpp -- device-side maths for the SPH solver core (gfx950).
//
// Own vector type and inline functions used by every kernel.  Expression
// order follows the reference kernels (kennychufk/alluvion runner.hpp) so that
// the FP32 results are reproducible; all code is compiled with
// -ffp-contract=off (no implicit fused multiply-add), and the square roots
// and divisions of the kernel functions (sqrt_rn, div_known, div_rn) are
// correctly rounded.
#pragma once

#include <float.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "alv_types.hpp"

namespace alv {

struct vec3 {
  float x, y, z;
};
__host__ __device__ inline vec3 mk(float x, float y, float z) {
  vec3 r;
  r.x = x;
  r.y = y;
  r.z = z;
  return r;
}
__host__ __device__ inline vec3 operator+(vec3 a, vec3 b) {
  return mk(a.x + b.x, a.y + b.y, a.z + b.z);
}
__host__ __device__ inline vec3 operator-(vec3 a, vec3 b) {
  return mk(a.x - b.x, a.y - b.y, a.z - b.z);
}
__host__ __device__ inline vec3 operator-(vec3 a) { return mk(-a.x, -a.y, -a.z); }
__host__ __device__ inline vec3 operator*(float s, vec3 a) {
  return mk(s * a.x, s * a.y, s * a.z);
}
__host__ __device__ inline vec3 operator*(vec3 a, float s) {
  return mk(a.x * s, a.y * s, a.z * s);
}
__host__ __device__ inline vec3 operator*(vec3 a, vec3 b) {
  return mk(a.x * b.x, a.y * b.y, a.z * b.z);
}
__host__ __device__ inline vec3& operator+=(vec3& a, vec3 b) {
  a = a + b;
  return a;
}
__host__ __device__ inline vec3& operator-=(vec3& a, vec3 b) {
  a = a - b;
  return a;
}
__host__ __device__ inline float dot(vec3 a, vec3 b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
__host__ __device__ inline vec3 cross(vec3 a, vec3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__host__ __device__ inline float length_sqr(vec3 a) { return dot(a, a); }

__device__ inline vec3 ld3(const float4& v) { return mk(v.x, v.y, v.z); }
__device__ inline vec3 ld3(const float* p, unsigned i) {
  return mk(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
}
__device__ inline void st3(float* p, unsigned i, vec3 v) {
  p[3 * i] = v.x;
  p[3 * i + 1] = v.y;
  p[3 * i + 2] = v.z;
}
__device__ inline float4 f4(vec3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

// Correctly rounded x / d for a divisor known in advance, from its correctly
// rounded reciprocal rd = RN(1/d) (computed once on the host): q = RN(x rd)
// is within one unit in the last place of x/d, the remainder x - q d is
// exact with one fused multiply-add, and RN(q + (x - q d) rd) = RN(x/d)
// (Markstein's theorem: P. Markstein, "IA-64 and Elementary Functions",
// 2000, section 8; N. Brisebarre, J.-M. Muller, S. K. Raina, "Accelerating
// correctly rounded floating-point division when the divisor is known in
// advance", IEEE Trans. Computers 53(8), 2004), for any x whose quotient
// neither overflows nor underflows.  Three instructions instead of the ten of
// the general correctly rounded division.  The kernel functions below divide
// by h and h^2 this way; their results are the bits of the IEEE divisions
// (a quotient below the normal range only occurs for particles closer than
// 1e-19 h, where W and gradW no longer depend on its last bits).
__device__ inline float div_known(float x, float d, float rd) {
  const float q = x * rd;
  const float r = __builtin_fmaf(-q, d, x);
  return __builtin_fmaf(r, rd, q);
}

// Square root rounded to nearest, from the hardware square root (within one
// unit in the last place, v_sqrt_f32): the estimate's neighbours s - 1 ulp
// and s + 1 ulp are tested with exact remainders x - s' s (fused
// multiply-add); a float x never lies within (s - ulp/2)^2 - ulp^2/4 .. of a
// midpoint square, so the tests pick RN(sqrt(x)).  Checked on the CPU for
// every float in [2^-96, 2^32] and every estimate within one ulp.  Smaller
// arguments (squared distances of particles closer than 1e-29 h) may round
// differently, but the kernel values below do not depend on those bits.
// Nine instructions instead of the fifteen of the general correctly rounded
// sqrtf, which also scales denormal arguments.
__device__ inline float sqrt_rn(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  const float s_dn = __uint_as_float(__float_as_uint(s) - 1u);
  const float s_up = __uint_as_float(__float_as_uint(s) + 1u);
  const float r_dn = __builtin_fmaf(-s_dn, s, x);
  const float r_up = __builtin_fmaf(-s_up, s, x);
  const float t = r_dn <= 0.0f ? s_dn : s;
  return r_up > 0.0f ? s_up : t;
}

// Reciprocal rounded to nearest for a normal y whose reciprocal is normal:
// one Newton step with an exact fused-multiply-add residual, from the
// hardware estimate (v_rcp_f32, within one ulp).  By Markstein's reciprocal
// theorem the step gives RN(1/y) from every estimate within one ulp, except
// for a significand of all ones with the estimate on the power of two below
// 1/y; checked on the CPU for every significand and every estimate within one
// ulp, and on the GPU for every significand (the estimate depends on the
// significand only).
__device__ inline float rcp_rn(float y) {
  const float r0 = __builtin_amdgcn_rcpf(y);
  const float e = __builtin_fmaf(-y, r0, 1.0f);
  return __builtin_fmaf(e, r0, r0);
}
// x / y rounded to nearest: the known-divisor sequence with y's correctly
// rounded reciprocal (six instructions instead of the general sequence's
// scaling and fix-up), for normal y and quotients in the normal range.
__device__ inline float div_rn(float x, float y) { return div_known(x, y, rcp_rn(y)); }

// --- cubic spline kernel (runner.hpp:596-638) ------------------------------
__device__ inline float W_dist(float r, const SimParams& p) {
  float q = div_known(fabsf(r), p.h, p.inv_h);
  float conj = 1.0f - q;
  float result = 0.0f;
  if (q <= 0.5f)
    result = (6.0f * q - 6.0f) * q * q + 1.0f;
  else if (q <= 1.0f)
    result = 2.0f * conj * conj * conj;
  return result * p.k;
}
// Both branch values are computed and one is selected, so the neighbour
// loops stay branch-free (the selected value is the same FP32 result).
__device__ inline float W_d2(float r2, const SimParams& p) {
  float q2 = div_known(r2, p.h2, p.inv_h2);
  float q = div_known(sqrt_rn(r2), p.h, p.inv_h);
  float conj = 1.0f - q;
  float inner = (6.0f * q - 6.0f) * q2 + 1.0f;
  float outer = 2.0f * conj * conj * conj;
  float result = q <= 0.5f ? inner : (q <= 1.0f ? outer : 0.0f);
  return result * p.k;
}
// gradW(d) = gradW_scale(d) * d.  The outer branch divides by r, which lies
// in (h/2, h] where it is selected (elsewhere its quotient is discarded):
// by div_rn, or by the compiler's general IEEE sequence (same bits).
template <bool kRn = true>
__device__ inline float gradW_scale(vec3 d, const SimParams& p) {
  float rl2 = length_sqr(d);
  float rl = sqrt_rn(rl2);
  float q = div_known(rl, p.h, p.inv_h);
  float s_in = div_known(q * 3.0f - 2.0f, p.h, p.inv_h);
  float s_out = kRn ? div_rn((1.0f - q) * (q - 1.0f), rl) : (1.0f - q) * (q - 1.0f) / rl;
  float scale = q <= 0.5f ? s_in : (q <= 1.0f ? s_out : 0.0f);
  return p.l * scale;
}
__device__ inline vec3 gradW(vec3 d, const SimParams& p) { return gradW_scale<false>(d, p) * d; }
__device__ inline vec3 gradW_rn(vec3 d, const SimParams& p) { return gradW_scale<true>(d, p) * d; }

// --- quaternion rotation (q = x, y, z, w) -----------------------------------
__host__ __device__ inline vec3 rotate(vec3 v, float4 q) {
  float xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
  float xy = q.x * q.y, xz = q.x * q.z, yz = q.y * q.z;
  float wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
  vec3 r;
  r.x = (1.0f - 2.0f * (yy + zz)) * v.x + 2.0f * (xy - wz) * v.y +
        2.0f * (xz + wy) * v.z;
  r.y = 2.0f * (xy + wz) * v.x + (1.0f - 2.0f * (xx + zz)) * v.y +
        2.0f * (yz - wx) * v.z;
  r.z = 2.0f * (xz - wy) * v.x + 2.0f * (yz + wx) * v.y +
        (1.0f - 2.0f * (xx + yy)) * v.z;
  return r;
}
__host__ __device__ inline float4 conjugate(float4 q) {
  return make_float4(-q.x, -q.y, -q.z, q.w);
}

// --- analytic SDFs (dg/box_distance.hpp, dg/box_shell_distance.hpp,
//     dg/sphere_distance.hpp; device variants) --------------------------------
__device__ inline float box_dist(vec3 x, vec3 half) {
  vec3 diff = mk(fabsf(x.x) - half.x, fabsf(x.y) - half.y, fabsf(x.z) - half.z);
  vec3 cl = mk(fmaxf(diff.x, 0.0f), fmaxf(diff.y, 0.0f), fmaxf(diff.z, 0.0f));
  return sqrtf(length_sqr(cl)) + fminf(fmaxf(diff.x, fmaxf(diff.y, diff.z)), 0.0f);
}
// Signed distance of an analytic boundary b at x (body frame, before the
// boundary's sign).  A distance map is sampled with map_distance instead:
// keeping its 32-node interpolation out of this function keeps it out of the
// register count of every kernel that evaluates an analytic distance.
__device__ inline float sdf_eval(const BoundaryDev& b, vec3 x) {
  if (b.type == ALV_DIST_BOX) return box_dist(x, mk(b.half[0], b.half[1], b.half[2])) - b.outset;
  if (b.type == ALV_DIST_BOX_SHELL) {
    float inner = box_dist(x, mk(b.half[0], b.half[1], b.half[2])) - b.outset;
    float outer =
        box_dist(x, mk(b.half_outer[0], b.half_outer[1], b.half_outer[2])) - b.outset;
    return fmaxf(outer, -inner);
  }
  return sqrtf(length_sqr(x)) - b.radius;
}
__device__ inline vec3 sdf_gradient(const BoundaryDev& b, vec3 x, float scale) {
  if (b.type == ALV_DIST_SPHERE) return x;
  float step = scale * 0.00390625f;
  return mk(sdf_eval(b, x + mk(step, 0.0f, 0.0f)) - sdf_eval(b, x - mk(step, 0.0f, 0.0f)),
            sdf_eval(b, x + mk(0.0f, step, 0.0f)) - sdf_eval(b, x - mk(0.0f, step, 0.0f)),
            sdf_eval(b, x + mk(0.0f, 0.0f, step)) - sdf_eval(b, x - mk(0.0f, 0.0f, step)));
}

// --- 32-node cubic Lagrange volume map (runner.hpp:1131-1594) ---------------
// Node numbering: (rx+1)(ry+1)(rz+1) corner nodes with x fastest, then two
// nodes on every x-, y- and z-directed cell edge.
__host__ __device__ inline vec3 node_position(const GridGeomDev& g, unsigned l) {
  const unsigned rx = g.res[0], ry = g.res[1], rz = g.res[2];
  unsigned nv = (rx + 1) * (ry + 1) * (rz + 1);
  unsigned ne_x = rx * (ry + 1) * (rz + 1);
  unsigned ne_y = (rx + 1) * ry * (rz + 1);
  vec3 dmin = mk(g.dmin[0], g.dmin[1], g.dmin[2]);
  vec3 cs = mk(g.cs[0], g.cs[1], g.cs[2]);
  vec3 x;
  if (l < nv) {
    unsigned temp = l % ((ry + 1) * (rx + 1));
    x = dmin + cs * mk(float(temp % (rx + 1)), float(temp / (rx + 1)),
                       float(l / ((ry + 1) * (rx + 1))));
  } else if (l < nv + 2 * ne_x) {
    l -= nv;
    unsigned e = l / 2;
    unsigned temp = e % ((ry + 1) * rx);
    x = dmin + cs * mk(float(temp % rx), float(temp / rx), float(e / ((ry + 1) * rx)));
    x.x += (1.0f + float(l % 2)) / 3.0f * cs.x;
  } else if (l < nv + 2 * (ne_x + ne_y)) {
    l -= (nv + 2 * ne_x);
    unsigned e = l / 2;
    unsigned temp = e % ((rz + 1) * ry);
    x = dmin + cs * mk(float(e / ((rz + 1) * ry)), float(temp % ry), float(temp / ry));
    x.y += (1.0f + float(l % 2)) / 3.0f * cs.y;
  } else {
    l -= (nv + 2 * (ne_x + ne_y));
    unsigned e = l / 2;
    unsigned temp = e % ((rx + 1) * rz);
    x = dmin + cs * mk(float(temp / rz), float(e / ((rx + 1) * rz)), float(temp % rz));
    x.z += (1.0f + float(l % 2)) / 3.0f * cs.z;
  }
  return x;
}

// Locates x in the grid; returns false outside the (closed) domain.
__device__ inline bool grid_resolve(const GridGeomDev& g, vec3 x, int ip[3], vec3& inner) {
  vec3 inv = mk(1.0f / g.cs[0], 1.0f / g.cs[1], 1.0f / g.cs[2]);
  if (x.x >= g.dmin[0] && x.y >= g.dmin[1] && x.z >= g.dmin[2] && g.dmax[0] >= x.x &&
      g.dmax[1] >= x.y && g.dmax[2] >= x.z) {
    vec3 f = (x - mk(g.dmin[0], g.dmin[1], g.dmin[2])) * inv;
    ip[0] = min(int(f.x), int(g.res[0]) - 1);
    ip[1] = min(int(f.y), int(g.res[1]) - 1);
    ip[2] = min(int(f.z), int(g.res[2]) - 1);
    vec3 sd_min = mk(g.dmin[0], g.dmin[1], g.dmin[2]) +
                  mk(g.cs[0], g.cs[1], g.cs[2]) * mk(float(ip[0]), float(ip[1]), float(ip[2]));
    vec3 t = (x - sd_min) * inv;
    inner = mk(2.0f * t.x - 1.0f, 2.0f * t.y - 1.0f, 2.0f * t.z - 1.0f);
    return true;
  }
  return false;
}

__host__ __device__ inline void grid_cell_nodes(const unsigned res[3], const int ip[3],
                                                unsigned* cell) {
  const unsigned rx = res[0], ry = res[1], rz = res[2];
  const unsigned i = ip[0], j = ip[1], k = ip[2];
  const unsigned sx = rx + 1, sxy = (rx + 1) * (ry + 1);
#pragma unroll
  for (unsigned c = 0; c < 8; ++c) {
    unsigned di = c & 1u, dj = (c >> 1) & 1u, dk = (c >> 2) & 1u;
    cell[c] = sxy * (k + dk) + sx * (j + dj) + (i + di);
  }
  unsigned base = (rx + 1) * (ry + 1) * (rz + 1);
#pragma unroll
  for (unsigned e = 0; e < 4; ++e) {
    unsigned dj = e >> 1, dk = e & 1u;
    unsigned n = base + 2 * (rx * (ry + 1) * (k + dk) + rx * (j + dj) + i);
    cell[8 + 2 * e] = n;
    cell[9 + 2 * e] = n + 1;
  }
  base += 2 * (rx * (ry + 1) * (rz + 1));
#pragma unroll
  for (unsigned e = 0; e < 4; ++e) {
    unsigned di = e & 1u, dk = e >> 1;
    unsigned n = base + 2 * (ry * (rz + 1) * (i + di) + ry * (k + dk) + j);
    cell[16 + 2 * e] = n;
    cell[17 + 2 * e] = n + 1;
  }
  base += 2 * ((rx + 1) * ry * (rz + 1));
#pragma unroll
  for (unsigned e = 0; e < 4; ++e) {
    unsigned di = e >> 1, dj = e & 1u;
    unsigned n = base + 2 * (rz * (rx + 1) * (j + dj) + rz * (i + di) + k);
    cell[24 + 2 * e] = n;
    cell[25 + 2 * e] = n + 1;
  }
}

// Value of the 32-node element at local coordinate xi, interpolating the
// nodal values `nodes` of the cell.  Corner node (a,b,c) in {-1,1}^3:
//   N = (1/64)(1+a x)(1+b y)(1+c z)(9(x^2+y^2+z^2) - 19);
// edge node at u = s/3 along an edge with transverse signs (p, q):
//   N = (9/64)(1-u^2)(1+3 s u)(1+p v1)(1+q v2).
__device__ inline float grid_interpolate(const float* __restrict__ nodes,
                                         const unsigned* cell, vec3 xi) {
  const float x = xi.x, y = xi.y, z = xi.z;
  const float k64 = 1.0f / 64.0f;
  const float k964 = 9.0f / 64.0f;
  const float F = k64 * (9.0f * (x * x + y * y + z * z) - 19.0f);
  float phi = 0.0f;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    float a = (c & 1) ? 1.0f : -1.0f;
    float b = (c & 2) ? 1.0f : -1.0f;
    float cc = (c & 4) ? 1.0f : -1.0f;
    float A = 1.0f + a * x, B = 1.0f + b * y, C = 1.0f + cc * z;
    float N = F * (A * B) * C;
    phi += nodes[cell[c]] * N;
  }
#pragma unroll
  for (int axis = 0; axis < 3; ++axis) {
    float u = axis == 0 ? x : (axis == 1 ? y : z);
    float v1 = axis == 0 ? y : x;
    float v2 = axis == 2 ? y : z;
    float one_m_u2 = 1.0f - u * u;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      float p, q;
      if (axis == 0) {
        p = (e >> 1) ? 1.0f : -1.0f;
        q = (e & 1) ? 1.0f : -1.0f;
      } else if (axis == 1) {
        p = (e & 1) ? 1.0f : -1.0f;
        q = (e >> 1) ? 1.0f : -1.0f;
      } else {
        p = (e >> 1) ? 1.0f : -1.0f;
        q = (e & 1) ? 1.0f : -1.0f;
      }
      float P = 1.0f + p * v1, Q = 1.0f + q * v2;
#pragma unroll
      for (int s2 = 0; s2 < 2; ++s2) {
        float s = s2 ? 1.0f : -1.0f;
        float L = 1.0f + 3.0f * s * u;
        int n = 8 + axis * 8 + 2 * e + s2;
        float G = k964 * (one_m_u2 * L);
        float N = G * (P * Q);
        phi += nodes[cell[n]] * N;
      }
    }
  }
  return phi;
}

// A distance map's distance and its gradient in world units (d/dx = (2 /
// cell size) d/dxi) at local coordinate xi of cell ip (interpolate_and_derive,
// runner.hpp:1773-1791) and, with Vol, the volume (grid_interpolate over
// `vol`), in one pass over the 32 nodes.  The shape functions' derivatives
// follow from the closed forms above:
//   corner: dN/dx = (1/64)(1+b y)(1+c z)[a (9 r^2 - 19) + 18 x (1+a x)];
//   edge along u: dN/du = (9/64)(1+p v1)(1+q v2)(3 s (1 - u^2) - 2 u (1+3 s u)),
//   dN/dv1 = (9/64)(1-u^2)(1+3 s u) p (1+q v2), likewise dN/dv2.
// The volume uses grid_interpolate's expressions and node order, so it is
// that function's value bit for bit.  The edge nodes are visited one axis at a
// time in a loop that is not unrolled, so at most 16 node loads are in flight
// rather than 64: fully unrolled, the interpolations set the register count
// (and so the occupancy) of every kernel that holds them, analytic
// boundaries included.
template <bool Vol>
__device__ inline float grid_map_eval(const float* __restrict__ dist,
                                      const float* __restrict__ vol, const unsigned res[3],
                                      const int ip[3], vec3 xi, const float cs[3], vec3& grad,
                                      float& volume) {
  const float x = xi.x, y = xi.y, z = xi.z;
  const float k64 = 1.0f / 64.0f;
  const float k964 = 9.0f / 64.0f;
  const float F = k64 * (9.0f * (x * x + y * y + z * z) - 19.0f);
  const unsigned rx = res[0], ry = res[1], rz = res[2];
  const unsigned i = ip[0], j = ip[1], k = ip[2];
  const unsigned sx = rx + 1, sxy = (rx + 1) * (ry + 1);
  float phi = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f, pv = 0.0f;
  {
    float dn[8], vn[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const unsigned n = sxy * (k + ((c >> 2) & 1)) + sx * (j + ((c >> 1) & 1)) + (i + (c & 1));
      dn[c] = dist[n];
      if (Vol) vn[c] = vol[n];
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      float a = (c & 1) ? 1.0f : -1.0f;
      float b = (c & 2) ? 1.0f : -1.0f;
      float cc = (c & 4) ? 1.0f : -1.0f;
      float A = 1.0f + a * x, B = 1.0f + b * y, C = 1.0f + cc * z;
      const float v = dn[c];
      phi += v * (F * (A * B) * C);
      gx += v * ((a * F + A * (k64 * (18.0f * x))) * (B * C));
      gy += v * ((b * F + B * (k64 * (18.0f * y))) * (A * C));
      gz += v * ((cc * F + C * (k64 * (18.0f * z))) * (A * B));
      if (Vol) pv += vn[c] * (F * (A * B) * C);
    }
  }
  // edge nodes of axis a: base_a + 2 (So (o + (e & 1)) + Sm (m + (e >> 1)) + c) + s
  const unsigned nv = sxy * (rz + 1);
  const unsigned ne_x = rx * (ry + 1) * (rz + 1), ne_y = (rx + 1) * ry * (rz + 1);
#pragma unroll 1
  for (int axis = 0; axis < 3; ++axis) {
    const float u = axis == 0 ? x : (axis == 1 ? y : z);
    const float v1 = axis == 0 ? y : x;
    const float v2 = axis == 2 ? y : z;
    const float one_m_u2 = 1.0f - u * u;
    unsigned base, So, Sm, o, m, c0;
    if (axis == 0) {
      base = nv, So = rx * (ry + 1), Sm = rx, o = k, m = j, c0 = i;
    } else if (axis == 1) {
      base = nv + 2 * ne_x, So = ry * (rz + 1), Sm = ry, o = i, m = k, c0 = j;
    } else {
      base = nv + 2 * (ne_x + ne_y), So = rz * (rx + 1), Sm = rz, o = j, m = i, c0 = k;
    }
    float dn[8], vn[8];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const unsigned n = base + 2 * (So * (o + (e & 1)) + Sm * (m + (e >> 1)) + c0);
      dn[2 * e] = dist[n];
      dn[2 * e + 1] = dist[n + 1];
      if (Vol) {
        vn[2 * e] = vol[n];
        vn[2 * e + 1] = vol[n + 1];
      }
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int pb = axis == 1 ? (e & 1) : (e >> 1), qb = axis == 1 ? (e >> 1) : (e & 1);
      const float pp = pb ? 1.0f : -1.0f, qq = qb ? 1.0f : -1.0f;
      float P = 1.0f + pp * v1, Q = 1.0f + qq * v2;
#pragma unroll
      for (int s2 = 0; s2 < 2; ++s2) {
        float sg = s2 ? 1.0f : -1.0f;
        float L = 1.0f + 3.0f * sg * u;
        const float v = dn[2 * e + s2];
        float G = k964 * (one_m_u2 * L);
        phi += v * (G * (P * Q));
        if (Vol) pv += vn[2 * e + s2] * (G * (P * Q));
        float du = k964 * (3.0f * sg * one_m_u2 - 2.0f * u * L) * (P * Q);
        float dv1 = G * (pp * Q);
        float dv2 = G * (P * qq);
        float dx = axis == 0 ? du : dv1;
        float dy = axis == 1 ? du : (axis == 0 ? dv1 : dv2);
        float dz = axis == 2 ? du : dv2;
        gx += v * dx;
        gy += v * dy;
        gz += v * dz;
      }
    }
  }
  grad = mk(gx * (2.0f / cs[0]), gy * (2.0f / cs[1]), gz * (2.0f / cs[2]));
  if (Vol) volume = pv;
  return phi;
}

// Interpolates the 32-node map at x; FLT_MAX outside (used by the volume
// field builder for distance maps, runner.hpp:1596-1614).
__device__ inline float grid_sample(const GridGeomDev& g, const float* __restrict__ nodes,
                                    vec3 x, float outside) {
  int ip[3];
  vec3 inner;
  if (!grid_resolve(g, x, ip, inner)) return outside;
  unsigned cell[32];
  grid_cell_nodes(g.res, ip, cell);
  return grid_interpolate(nodes, cell, inner);
}

__device__ inline float map_distance(const BoundaryDev& b, vec3 x) {
  return grid_sample(b.geom, b.distance, x, FLT_MAX);
}

// --- neighbour grid helpers (runner.hpp:1827-1838, 689-707) -----------------
// Boundary record of point x for boundary B
// (compute_volume_and_boundary_x_analytic, runner.hpp:1738-1771, and
// compute_particle_boundary_analytic, runner.hpp:1982-2002; for a distance
// map compute_volume_and_boundary_x, runner.hpp:1696-1736, and
// compute_particle_boundary, runner.hpp:2005-2025): rec = (closest boundary
// point, volume Psi), ker = (Psi gradW, Psi W).  Returns Psi.
// With Sparse, a point outside the map (first by a sphere around the body
// origin that holds the map's box, before the rotation into the body frame)
// or in a cell whose nodes are all zero returns 0 at once and leaves rec and
// ker unset (its volume is exactly 0, so every term of the record is too).
template <bool Sparse>
__device__ inline float boundary_record(const BoundaryDev& B, vec3 x, const SimParams& p,
                                        float4& rec, float4& ker) {
  const float4 q = make_float4(B.q[0], B.q[1], B.q[2], B.q[3]);
  const vec3 rel = x - mk(B.x[0], B.x[1], B.x[2]);
  // farther from the body origin than any point of the map: outside it
  if (Sparse && length_sqr(rel) > B.map_radius2) return 0.0f;
  vec3 local = rotate(rel, conjugate(q));
  int ip[3];
  vec3 inner;
  const bool in_map = grid_resolve(B.geom, local, ip, inner);
  if constexpr (Sparse) {
    if (!in_map) return 0.0f;
    if (B.cell_nz &&
        !B.cell_nz[(unsigned(ip[2]) * B.geom.res[1] + unsigned(ip[1])) * B.geom.res[0] + unsigned(ip[0])])
      return 0.0f;
  }
  float d;
  vec3 normal;
  float vol = 0.0f;
  if (B.type == ALV_DIST_MAP) {
    // compute_volume_and_boundary_x (runner.hpp:1696-1736): distance and
    // normal from the distance map; outside the map the record is x itself
    // with zero volume
    d = 0.0f;
    normal = mk(0.f, 0.f, 0.f);
    if (in_map) {
      vec3 g;
      d = grid_map_eval<true>(B.distance, B.volume, B.geom.res, ip, inner, B.geom.cs, g, vol) *
          B.sign;
      normal = rotate(g * B.sign, q);
    }
  } else {
    d = sdf_eval(B, local) * B.sign;
    normal = rotate(sdf_gradient(B, local, p.h) * B.sign, q);
    if (in_map) {
      unsigned cell[32];
      grid_cell_nodes(B.geom.res, ip, cell);
      vol = grid_interpolate(B.volume, cell, inner);
    }
  }
  float nl2 = length_sqr(normal);
  normal = normal * (nl2 > 0.0f ? 1.0f / sqrtf(nl2) : 0.0f);
  vec3 xi_bxj = d * normal;
  vec3 bx = x - xi_bxj;
  vec3 kg = gradW(xi_bxj, p);
  rec = make_float4(bx.x, bx.y, bx.z, vol);
  ker = make_float4(vol * kg.x, vol * kg.y, vol * kg.z, vol * W_dist(d, p));
  return vol;
}

__device__ inline void cell_ipos(vec3 x, const SimParams& p, int ip[3]) {
  ip[0] = int(floorf(x.x / p.h)) - p.off[0];
  ip[1] = int(floorf(x.y / p.h)) - p.off[1];
  ip[2] = int(floorf(x.z / p.h)) - p.off[2];
}
__device__ inline bool within_grid(const int ip[3], const SimParams& p) {
  return 0 <= ip[0] && ip[0] < p.res[0] && 0 <= ip[1] && ip[1] < p.res[1] && 0 <= ip[2] &&
         ip[2] < p.res[2];
}
__device__ inline unsigned cell_linear(const int ip[3], const SimParams& p) {
  return (unsigned(ip[0]) * unsigned(p.res[1]) + unsigned(ip[1])) * unsigned(p.res[2]) +
         unsigned(ip[2]);
}
__device__ inline float wrap1(float d, float len, float lo, float hi) {
  if (d >= hi)
    d -= len;
  else if (d < lo)
    d += len;
  return d;
}
__device__ inline vec3 wrap_disp(vec3 d, const SimParams& p) {
  if (p.wrap_mask & 1) d.x = wrap1(d.x, p.wrap_len[0], p.wrap_min[0], p.wrap_max[0]);
  if (p.wrap_mask & 2) d.y = wrap1(d.y, p.wrap_len[1], p.wrap_min[1], p.wrap_max[1]);
  if (p.wrap_mask & 4) d.z = wrap1(d.z, p.wrap_len[2], p.wrap_min[2], p.wrap_max[2]);
  return d;
}

// Block index remapped so that each of the 8 XCDs (which receive blocks
// round-robin) works on one contiguous range of particles: neighbouring
// particles then share an XCD's L2.
__device__ inline unsigned xcd_block() {
  const unsigned b = blockIdx.x, G = gridDim.x;
  const unsigned q = G / 8u, r = G % 8u;
  const unsigned x = b % 8u, k = b / 8u;
  return x * q + min(x, r) + k;
}

}  // namespace alv
