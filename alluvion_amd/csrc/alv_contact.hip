// alv_contact.hip -- rigid-body contact detection (reference runner.hpp:3833-
// 3914 `collision_test`, analytic-distance variant; pile.hpp:826-904
// `find_contacts`).
//
// Every collision vertex of body i is tested against the signed distance of
// body j (box or sphere, in j's frame).  The reference appends contacts with
// an atomic counter, so their order -- and the result of the sequential
// Gauss-Seidel contact solve that follows -- depends on thread timing.  Here
// the candidates are numbered (pair in (i, j) order, then vertex) and one
// workgroup compacts them with a wave-ballot scan, so contacts come out in
// that order on every run.  Contact counts are small (the vertices of a few
// collision meshes), so a single workgroup is enough; the kernel is far from
// any hardware bound and is not on the fluid's hot path.
#include <hip/hip_runtime.h>

#include "alv_contact.hpp"
#include "alv_device.hpp"

namespace alv {
namespace {

constexpr unsigned kContactThreads = 1024;
constexpr unsigned kContactWaves = kContactThreads / 64;

// R diag(v) R^T for the rotation of q, stored as diagonal and (01, 02, 12)
// off-diagonal (runner.hpp:178-199).
__device__ inline void congruent_matrix(vec3 v, float4 q, vec3* diag, vec3* off) {
  const float m00 = 1 - 2 * (q.y * q.y + q.z * q.z), m01 = 2 * (q.x * q.y - q.z * q.w),
              m02 = 2 * (q.x * q.z + q.y * q.w);
  const float m10 = 2 * (q.x * q.y + q.z * q.w), m11 = 1 - 2 * (q.x * q.x + q.z * q.z),
              m12 = 2 * (q.y * q.z - q.x * q.w);
  const float m20 = 2 * (q.x * q.z - q.y * q.w), m21 = 2 * (q.y * q.z + q.x * q.w),
              m22 = 1 - 2 * (q.x * q.x + q.y * q.y);
  *diag = mk(v.x * m00 * m00 + v.y * m01 * m01 + v.z * m02 * m02,
             v.x * m10 * m10 + v.y * m11 * m11 + v.z * m12 * m12,
             v.x * m20 * m20 + v.y * m21 * m21 + v.z * m22 * m22);
  *off = mk(v.x * m00 * m10 + v.y * m01 * m11 + v.z * m02 * m12,
            v.x * m00 * m20 + v.y * m01 * m21 + v.z * m02 * m22,
            v.x * m10 * m20 + v.y * m11 * m21 + v.z * m12 * m22);
}

// Impulse response matrix K = I/m - [r]x I_w^-1 [r]x of one body at lever
// arm r (runner.hpp:265-288); zero for a body of zero mass.
__device__ inline void contact_k(vec3 r, float mass, vec3 d, vec3 o, vec3* kd, vec3* ko) {
  *kd = mk(0.f, 0.f, 0.f);
  *ko = mk(0.f, 0.f, 0.f);
  if (mass == 0.f) return;
  const float im = 1 / mass;
  *kd = mk(r.z * r.z * d.y - r.y * r.z * o.z * 2 + r.y * r.y * d.z + im,
           r.z * r.z * d.x - r.x * r.z * o.y * 2 + r.x * r.x * d.z + im,
           r.y * r.y * d.x - r.x * r.y * o.x * 2 + r.x * r.x * d.y + im);
  *ko = mk(-r.z * r.z * o.x + r.x * r.z * o.z + r.y * r.z * o.y - r.x * r.y * d.z,
           r.y * r.z * o.x - r.x * r.z * d.y - r.y * r.y * o.y + r.x * r.y * o.z,
           -(r.y * r.z * d.x) + r.x * r.z * o.x + r.x * r.y * o.y - r.x * r.x * o.z);
}

__device__ inline vec3 sym_apply(vec3 v, vec3 d, vec3 o) {
  return mk(v.x * d.x + v.y * o.x + v.z * o.y, v.x * o.x + v.y * d.y + v.z * o.z,
            v.x * o.y + v.y * o.z + v.z * d.z);
}

__device__ inline void put3(float* dst, vec3 v) {
  dst[0] = v.x;
  dst[1] = v.y;
  dst[2] = v.z;
}

// Tests vertex `vert` of body i against body j; true and a filled record
// when the vertex is inside body j (dist < 0).  World vertices (pellets,
// collision_test_with_pellets, runner.hpp:3627-3727) are already placed;
// body-frame vertices are placed with body i's kinematics.  `inertia`
// holds [B][3] principal inertia, then [B][2] (restitution, friction).
__device__ bool test_vertex(unsigned i, unsigned j, unsigned flags, unsigned nb,
                            const BoundaryDev* bodies, const float* inertia, const float* verts,
                            unsigned vert, float tolerance, alv_contact* out) {
  const BoundaryDev& Bi = bodies[i];
  const BoundaryDev& Bj = bodies[j];
  if ((flags & kJobSkip) && (i == j || (Bi.mass == 0.f && Bj.mass == 0.f))) return false;
  const float4 qi = make_float4(Bi.q[0], Bi.q[1], Bi.q[2], Bi.q[3]);
  const float4 qj = make_float4(Bj.q[0], Bj.q[1], Bj.q[2], Bj.q[3]);
  const vec3 xi = mk(Bi.x[0], Bi.x[1], Bi.x[2]), xj = mk(Bj.x[0], Bj.x[1], Bj.x[2]);
  const vec3 vx = mk(verts[3 * vert], verts[3 * vert + 1], verts[3 * vert + 2]);
  const vec3 cp_i = (flags & kJobWorld) ? vx : rotate(vx, qi) + xi;
  const vec3 wrt_j = rotate(cp_i - xj, conjugate(qj));
  float dist;
  vec3 n;
  if (Bj.type == ALV_DIST_MAP) {
    // collision_find_dist_normal (runner.hpp:1794-1825): outside the map
    // only the sign is known (no contact for a positive sign)
    dist = Bj.sign;
    n = mk(0.f, 0.f, 0.f);
    int ip[3];
    vec3 inner;
    if (grid_resolve(Bj.geom, wrt_j, ip, inner)) {
      vec3 g;
      float unused;
      dist = Bj.sign * grid_map_eval<false>(Bj.distance, nullptr, Bj.geom.res, ip, inner,
                                            Bj.geom.cs, g, unused) -
             tolerance;
      n = g * Bj.sign;
    }
  } else {
    dist = sdf_eval(Bj, wrt_j) * Bj.sign - tolerance;
    n = sdf_gradient(Bj, wrt_j, 1e-3f) * Bj.sign;
  }
  if (!(dist < 0)) return false;
  n = n * (1.0f / sqrtf(length_sqr(n)));
  const vec3 cp_j = rotate(wrt_j - dist * n, qj) + xj;
  const vec3 nw = rotate(n, qj);
  const vec3 r_i = cp_i - xi, r_j = cp_j - xj;
  const vec3 u_i = mk(Bi.v[0], Bi.v[1], Bi.v[2]) + cross(mk(Bi.omega[0], Bi.omega[1], Bi.omega[2]), r_i);
  const vec3 u_j = mk(Bj.v[0], Bj.v[1], Bj.v[2]) + cross(mk(Bj.omega[0], Bj.omega[1], Bj.omega[2]), r_j);
  const vec3 u_rel = u_i - u_j;
  const float u_rel_n = dot(nw, u_rel);
  vec3 t = u_rel - u_rel_n * nw;
  if (length_sqr(t) > 1e-6f) t = t * (1.0f / sqrtf(length_sqr(t)));
  const float* Ii = inertia + 3 * i;
  const float* Ij = inertia + 3 * j;
  const float* rf = inertia + 3 * nb;
  vec3 di, oi, dj, oj;
  congruent_matrix(mk(1 / Ii[0], 1 / Ii[1], 1 / Ii[2]), qi, &di, &oi);
  congruent_matrix(mk(1 / Ij[0], 1 / Ij[1], 1 / Ij[2]), qj, &dj, &oj);
  vec3 kdi, koi, kdj, koj;
  contact_k(r_i, Bi.mass, di, oi, &kdi, &koi);
  contact_k(r_j, Bj.mass, dj, oj, &kdj, &koj);
  const vec3 kd = kdi + kdj, ko = koi + koj;
  out->i = i;
  out->j = j;
  put3(out->cp_i, cp_i);
  put3(out->cp_j, cp_j);
  put3(out->n, nw);
  put3(out->t, t);
  put3(out->iiwi_diag, di);
  put3(out->iiwi_off_diag, oi);
  put3(out->iiwj_diag, dj);
  put3(out->iiwj_off_diag, oj);
  out->friction = rf[2 * i + 1] + rf[2 * j + 1];
  out->nkninv = 1 / dot(nw, sym_apply(nw, kd, ko));
  out->pmax = 1 / dot(t, sym_apply(t, kd, ko)) * dot(u_rel, t);
  out->goalu = u_rel_n < 0 ? -(rf[2 * i] * rf[2 * j]) * u_rel_n : 0.0f;
  out->impulse_sum = 0;
  return true;
}

__global__ __launch_bounds__(kContactThreads) void k_collision(
    const ContactJob* jobs, unsigned num_jobs, unsigned num_candidates, unsigned nb,
    const BoundaryDev* bodies, const float* inertia, const float* verts, float tolerance,
    unsigned max_contacts, alv_contact* contacts, unsigned* count) {
  __shared__ unsigned wave_sum[kContactWaves];
  __shared__ unsigned base;
  const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) base = 0;
  __syncthreads();
  for (unsigned c0 = 0; c0 < num_candidates; c0 += kContactThreads) {
    const unsigned c = c0 + threadIdx.x;
    alv_contact rec;
    bool hit = false;
    if (c < num_candidates) {
      unsigned k = 0;
      while (k + 1 < num_jobs && jobs[k + 1].cbase <= c) ++k;
      const ContactJob J = jobs[k];
      const unsigned local = c - J.cbase;
      hit = test_vertex(J.i, J.j + local % J.nj, J.flags, nb, bodies, inertia, verts,
                        J.vbase + local / J.nj, tolerance, &rec);
    }
    const unsigned long long mask = __ballot(hit);
    const unsigned before = __popcll(mask & ((1ull << lane) - 1ull));
    if (lane == 0) wave_sum[wave] = unsigned(__popcll(mask));
    __syncthreads();
    unsigned offset = base;
    for (unsigned w = 0; w < wave; ++w) offset += wave_sum[w];
    const unsigned slot = offset + before;
    if (hit && slot < max_contacts) contacts[slot] = rec;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned total = 0;
      for (unsigned w = 0; w < kContactWaves; ++w) total += wave_sum[w];
      base += total;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) count[0] = base;
}

}  // namespace

void launch_collision(const ContactJob* jobs, unsigned num_jobs, unsigned num_candidates,
                      unsigned nb, const BoundaryDev* bodies, const float* inertia,
                      const float* verts, float tolerance, unsigned max_contacts,
                      alv_contact* contacts, unsigned* count, hipStream_t s) {
  hipLaunchKernelGGL(k_collision, dim3(1), dim3(kContactThreads), 0, s, jobs, num_jobs,
                     num_candidates, nb, bodies, inertia, verts, tolerance, max_contacts,
                     contacts, count);
}

}  // namespace alv
