// alv_sample.hip -- sampling the fluid at arbitrary points and the 256-bin
// histogram / KL divergence used to compare sampled fields
// (runner.hpp:1869-1911 make_neighbor_list for samples, 4173-4356 sample
// kernels, 4425-4537 histogram256 / quantize, 746-762 KLDivergence).
//
// Sample neighbour lists have the reference's format: [num_samples][K]
// float4 entries (x_sample - x_j, particle id j in the bits of w).  They are
// built against the grid of the solver's last neighbour search, with the
// particle positions of that search, as the reference's pid grid holds them.
#include <hip/hip_runtime.h>

#include "alv_device.hpp"
#include "alv_kernels.hpp"

namespace alv {
namespace {

constexpr unsigned kSampleBlock = 256;

inline unsigned sample_grid(unsigned n) { return (n + kSampleBlock - 1) / kSampleBlock; }

__device__ inline vec3 ld3f(const float* a, unsigned i) { return mk(a[3 * i], a[3 * i + 1], a[3 * i + 2]); }

// make_neighbor_list<wrap> on sample points.  As in the reference, a
// particle whose id equals the sample's index is not taken as its neighbour.
// With pellets (pcell_start non-null) each cell's pellets follow its fluid
// particles, as ids max_n + k follow every fluid id in the reference's grid.
template <bool W>
__global__ __launch_bounds__(kSampleBlock) void k_sample_neighbors(
    SimParams p, const unsigned* __restrict__ cell_start, const float4* __restrict__ xs,
    const unsigned* __restrict__ sorted_id, const unsigned* __restrict__ pcell_start,
    const float4* __restrict__ pxs, const unsigned* __restrict__ psorted_id, unsigned pid_base,
    const float* __restrict__ sample_x, unsigned ns, float4* __restrict__ out,
    unsigned* __restrict__ counts, unsigned* anomaly) {
  const unsigned i = blockIdx.x * kSampleBlock + threadIdx.x;
  if (i >= ns) return;
  const unsigned K = p.kmax;
  const vec3 x = ld3f(sample_x, i);
  int ip[3];
  cell_ipos(x, p, ip);
  unsigned cnt = 0;
  auto visit = [&](unsigned id, vec3 pos) {
    const vec3 d = W ? wrap_disp(x - pos, p) : x - pos;
    if (id != i && length_sqr(d) < p.h2) {
      out[size_t(i) * K + min(cnt, K - 1u)] = make_float4(d.x, d.y, d.z, __uint_as_float(id));
      ++cnt;
    }
  };
  for (int c = 0; c < 27; ++c) {
    int n[3] = {ip[0] + c / 9 - 1, ip[1] + (c / 3) % 3 - 1, ip[2] + c % 3 - 1};
#pragma unroll
    for (int a = 0; a < 3; ++a)
      if (p.wrap_mask & (1 << a)) n[a] = n[a] >= p.res[a] ? n[a] - p.res[a] : (n[a] < 0 ? n[a] + p.res[a] : n[a]);
    if (!within_grid(n, p)) continue;
    const unsigned cell = p.cell_base + cell_linear(n, p);
    for (unsigned j = cell_start[cell], je = cell_start[cell + 1]; j < je; ++j)
      visit(sorted_id[j], ld3(xs[j]));
    if (pcell_start)
      for (unsigned j = pcell_start[cell], je = pcell_start[cell + 1]; j < je; ++j)
        visit(pid_base + psorted_id[j], ld3(pxs[j]));
  }
  counts[i] = min(cnt, K);
  if (cnt > K) anomaly[2] = 1u;
}

// make_bead_pellet_neighbor_list<wrap> (runner.hpp:1913-1971): samples
// offset .. offset+ns of sample_x, their neighbours split by id into beads
// (id < max_beads) and pellets, rows indexed by the sample's particle index.
template <bool W>
__global__ __launch_bounds__(kSampleBlock) void k_sample_bead_pellet(
    SimParams p, const unsigned* __restrict__ cell_start, const float4* __restrict__ xs,
    const unsigned* __restrict__ sorted_id, const unsigned* __restrict__ pcell_start,
    const float4* __restrict__ pxs, const unsigned* __restrict__ psorted_id, unsigned pid_base,
    const float* __restrict__ sample_x, unsigned ns, unsigned offset, unsigned max_beads,
    float4* __restrict__ bead, unsigned* __restrict__ bead_counts, float4* __restrict__ pel,
    unsigned* __restrict__ pel_counts, unsigned* anomaly) {
  const unsigned i0 = blockIdx.x * kSampleBlock + threadIdx.x;
  if (i0 >= ns) return;
  const unsigned i = i0 + offset;
  const unsigned K = p.kmax;
  const vec3 x = ld3f(sample_x, i);
  int ip[3];
  cell_ipos(x, p, ip);
  unsigned cb = 0, cp = 0;
  auto visit = [&](unsigned id, vec3 pos) {
    const vec3 d = W ? wrap_disp(x - pos, p) : x - pos;
    if (id != i && length_sqr(d) < p.h2) {
      const float4 e = make_float4(d.x, d.y, d.z, __uint_as_float(id));
      if (id < max_beads)
        bead[size_t(i) * K + min(cb++, K - 1u)] = e;
      else
        pel[size_t(i) * K + min(cp++, K - 1u)] = e;
    }
  };
  for (int c = 0; c < 27; ++c) {
    int n[3] = {ip[0] + c / 9 - 1, ip[1] + (c / 3) % 3 - 1, ip[2] + c % 3 - 1};
#pragma unroll
    for (int a = 0; a < 3; ++a)
      if (p.wrap_mask & (1 << a)) n[a] = n[a] >= p.res[a] ? n[a] - p.res[a] : (n[a] < 0 ? n[a] + p.res[a] : n[a]);
    if (!within_grid(n, p)) continue;
    const unsigned cell = p.cell_base + cell_linear(n, p);
    for (unsigned j = cell_start[cell], je = cell_start[cell + 1]; j < je; ++j)
      visit(sorted_id[j], ld3(xs[j]));
    if (pcell_start)
      for (unsigned j = pcell_start[cell], je = pcell_start[cell + 1]; j < je; ++j)
        visit(pid_base + psorted_id[j], ld3(pxs[j]));
  }
  bead_counts[i] = min(cb, K);
  pel_counts[i] = min(cp, K);
  if (cb > K || cp > K) anomaly[2] = 1u;
}

// Sampling with pellet lists (runner.hpp:2122-2147 sample_position_density_
// with_pellets, 4242-4277 compute_sample_velocity_with_pellets, 4325-4359
// compute_sample_vorticity_with_pellets): fluid entries weighted m / rho_j,
// pellet entries by the particle volume; velocities of both from the user
// order rows of v (pellets at rows max_n + k).  what 0: density, 1:
// velocity, 2: vorticity (sample_v is the samples' velocity).
__global__ __launch_bounds__(kSampleBlock) void k_sample_with_pellets(
    SimParams p, int what, const float4* __restrict__ nbr, const unsigned* __restrict__ counts,
    const float4* __restrict__ pel, const unsigned* __restrict__ pcounts, unsigned K,
    unsigned ns, const float* __restrict__ density, const float* __restrict__ v,
    const float* __restrict__ sample_v, float* __restrict__ out) {
  const unsigned i = blockIdx.x * kSampleBlock + threadIdx.x;
  if (i >= ns) return;
  const unsigned nb = min(counts[i], K), np_ = min(pcounts[i], K);
  if (what == 0) {
    float dens = 0.f;
    for (unsigned t = 0; t < nb; ++t) dens += W_d2(length_sqr(ld3(nbr[size_t(i) * K + t])), p);
    for (unsigned t = 0; t < np_; ++t) dens += W_d2(length_sqr(ld3(pel[size_t(i) * K + t])), p);
    out[i] = dens * p.mass;
    return;
  }
  const vec3 vi = what == 2 ? ld3f(sample_v, i) : mk(0.f, 0.f, 0.f);
  vec3 r = mk(0.f, 0.f, 0.f);
  for (unsigned t = 0; t < nb; ++t) {
    const float4 e = nbr[size_t(i) * K + t];
    const unsigned j = __float_as_uint(e.w);
    const vec3 d = ld3(e);
    if (what == 1)
      r += (p.mass / density[j] * W_d2(length_sqr(d), p)) * ld3f(v, j);
    else
      r += (p.mass / density[j]) * cross(ld3f(v, j) - vi, gradW(d, p));
  }
  for (unsigned t = 0; t < np_; ++t) {
    const float4 e = pel[size_t(i) * K + t];
    const unsigned j = __float_as_uint(e.w);
    const vec3 d = ld3(e);
    if (what == 1)
      r += (p.vol * W_d2(length_sqr(d), p)) * ld3f(v, j);
    else
      r += p.vol * cross(ld3f(v, j) - vi, gradW(d, p));
  }
  out[3 * i] = r.x;
  out[3 * i + 1] = r.y;
  out[3 * i + 2] = r.z;
}

// Kernel sums over reference-layout lists (xixj, id) of particles with
// pellets.  compute_particle_boundary_with_pellets (runner.hpp:2027-2050):
// vol (sum grad W, sum W) over the pellet list;
// compute_density_with_pellets (runner.hpp:2074-2098): m (W(0) + sum W +
// sum W) over both lists; compute_dfsph_factor_with_pellets
// (runner.hpp:3156-3190): the DFSPH factor with the pellet list's entries
// taken as the boundary's (vol grad W) vectors, as the reference reads them.
enum { kPelletBoundary = 0, kPelletDensity = 1, kPelletFactor = 2 };
__global__ __launch_bounds__(kSampleBlock) void k_pellet_sums(
    SimParams p, int what, const float4* __restrict__ bead, const unsigned* __restrict__ bead_counts,
    const float4* __restrict__ pel, const unsigned* __restrict__ pel_counts, unsigned K,
    unsigned n, float* __restrict__ out) {
  const unsigned i = blockIdx.x * kSampleBlock + threadIdx.x;
  if (i >= n) return;
  const unsigned nb = bead ? min(bead_counts[i], K) : 0u;
  const unsigned np_ = min(pel_counts[i], K);
  if (what == kPelletBoundary) {
    vec3 g = mk(0.f, 0.f, 0.f);
    float w = 0.f;
    for (unsigned t = 0; t < np_; ++t) {
      const vec3 d = ld3(pel[size_t(i) * K + t]);
      w += W_d2(length_sqr(d), p);
      g += gradW(d, p);
    }
    reinterpret_cast<float4*>(out)[i] = make_float4(p.vol * g.x, p.vol * g.y, p.vol * g.z, p.vol * w);
  } else if (what == kPelletDensity) {
    float dens = p.w0;
    for (unsigned t = 0; t < nb; ++t) dens += W_d2(length_sqr(ld3(bead[size_t(i) * K + t])), p);
    for (unsigned t = 0; t < np_; ++t) dens += W_d2(length_sqr(ld3(pel[size_t(i) * K + t])), p);
    out[i] = dens * p.mass;
  } else {
    float sgk = 0.f;
    vec3 gpi = mk(0.f, 0.f, 0.f);
    for (unsigned t = 0; t < nb; ++t) {
      const vec3 gp = (-p.vol) * gradW(ld3(bead[size_t(i) * K + t]), p);
      sgk += length_sqr(gp);
      gpi -= gp;
    }
    for (unsigned t = 0; t < np_; ++t) {
      const vec3 gp = -ld3(pel[size_t(i) * K + t]);
      sgk += length_sqr(gp);
      gpi -= gp;
    }
    sgk += length_sqr(gpi);
    out[i] = sgk > p.eps ? -1.0f / sgk : 0.0f;
  }
}

// boundary records of sample points for every boundary, written in full
// ([B][num_samples], zero volume included), as compute_particle_boundary
// does for its points
__global__ __launch_bounds__(kSampleBlock) void k_sample_boundary(
    SimParams p, const BoundaryDev* __restrict__ bodies, unsigned nb,
    const float* __restrict__ sample_x, unsigned ns, float4* __restrict__ rec,
    float4* __restrict__ ker) {
  const unsigned i = blockIdx.x * kSampleBlock + threadIdx.x;
  if (i >= ns) return;
  const vec3 x = ld3f(sample_x, i);
  for (unsigned b = 0; b < nb; ++b) {
    float4 r, k;
    boundary_record<false>(bodies[b], x, p, r, k);
    rec[size_t(b) * ns + i] = r;
    ker[size_t(b) * ns + i] = k;
  }
}

// sample_fluid (runner.hpp:4173-4192): sum_j m / rho_j W q_j, q scalar or
// 3-vector
__global__ __launch_bounds__(kSampleBlock) void k_sample_fluid(
    SimParams p, const float4* __restrict__ nbr, const unsigned* __restrict__ counts, unsigned K,
    unsigned ns, const float* __restrict__ density, const float* __restrict__ q, unsigned comps,
    float* __restrict__ out) {
  const unsigned i = blockIdx.x * kSampleBlock + threadIdx.x;
  if (i >= ns) return;
  float acc[3] = {0.f, 0.f, 0.f};
  const unsigned nn = min(counts[i], K);
  for (unsigned k = 0; k < nn; ++k) {
    const float4 e = nbr[size_t(i) * K + k];
    const unsigned j = __float_as_uint(e.w);
    const float w = p.mass / density[j] * W_d2(e.x * e.x + e.y * e.y + e.z * e.z, p);
    for (unsigned c = 0; c < comps; ++c) acc[c] += w * q[size_t(j) * comps + c];
  }
  for (unsigned c = 0; c < comps; ++c) out[size_t(i) * comps + c] = acc[c];
}

// sample_fluid_density (runner.hpp:4193-4209): rho0 sum_j V W; with the
// sample boundary kernels, sample_position_density (runner.hpp:2100-2120)
// adds sum_b Psi_b W_b
__global__ __launch_bounds__(kSampleBlock) void k_sample_density(
    SimParams p, const float4* __restrict__ nbr, const unsigned* __restrict__ counts, unsigned K,
    unsigned ns, const float4* __restrict__ ker, unsigned nb, float* __restrict__ out) {
  const unsigned i = blockIdx.x * kSampleBlock + threadIdx.x;
  if (i >= ns) return;
  float dens = 0.f;
  const unsigned nn = min(counts[i], K);
  for (unsigned k = 0; k < nn; ++k) {
    const float4 e = nbr[size_t(i) * K + k];
    dens += p.vol * W_d2(e.x * e.x + e.y * e.y + e.z * e.z, p);
  }
  for (unsigned b = 0; b < nb; ++b) dens += ker[size_t(b) * ns + i].w;
  out[i] = dens * p.rho0;
}

__device__ inline vec3 body_velocity(const float* rx, const float* rv, const float* rw,
                                     unsigned b, vec3 bx) {
  return cross(ld3f(rw, b), bx - ld3f(rx, b)) + ld3f(rv, b);
}

// compute_sample_velocity (runner.hpp:4211-4241) and, with sample_v given,
// compute_sample_vorticity (runner.hpp:4289-4324)
template <bool Vort>
__global__ __launch_bounds__(kSampleBlock) void k_sample_velocity(
    SimParams p, const float4* __restrict__ nbr, const unsigned* __restrict__ counts, unsigned K,
    unsigned ns, const float* __restrict__ density, const float* __restrict__ v,
    const float* __restrict__ sample_v, const float4* __restrict__ rec,
    const float4* __restrict__ ker, unsigned nb, const float* __restrict__ rx,
    const float* __restrict__ rv, const float* __restrict__ rw, float* __restrict__ out) {
  const unsigned i = blockIdx.x * kSampleBlock + threadIdx.x;
  if (i >= ns) return;
  vec3 acc = mk(0.f, 0.f, 0.f);
  const vec3 vi = Vort ? ld3f(sample_v, i) : mk(0.f, 0.f, 0.f);
  const unsigned nn = min(counts[i], K);
  for (unsigned k = 0; k < nn; ++k) {
    const float4 e = nbr[size_t(i) * K + k];
    const unsigned j = __float_as_uint(e.w);
    const vec3 d = mk(e.x, e.y, e.z);
    if (Vort)
      acc += p.mass / density[j] * cross(ld3f(v, j) - vi, gradW(d, p));
    else
      acc += p.mass / density[j] * W_d2(length_sqr(d), p) * ld3f(v, j);
  }
  for (unsigned b = 0; b < nb; ++b) {
    const float4 r = rec[size_t(b) * ns + i], kk = ker[size_t(b) * ns + i];
    const vec3 velj = body_velocity(rx, rv, rw, b, mk(r.x, r.y, r.z));
    if (Vort)
      acc += cross(velj - vi, mk(kk.x, kk.y, kk.z));
    else
      acc += velj * kk.w;
  }
  out[3 * i] = acc.x;
  out[3 * i + 1] = acc.y;
  out[3 * i + 2] = acc.z;
}

// 256-bin histogram of floor((v - v_min) * step_inverse) mod 256 (quantize,
// runner.hpp:4517-4539: a negative or NaN argument gives bin 0, one beyond
// the 32-bit range bin 255).  One LDS histogram per wave, integer atomics, so
// the counts do not depend on the order of the additions.
__device__ inline unsigned quantize_bin(float v, float v_min, float step_inverse) {
  const float x = (v - v_min) * step_inverse;
  unsigned q;
  if (!(x >= 0.0f))
    q = 0u;
  else if (x >= 4294967296.0f)
    q = 0xFFFFFFFFu;
  else
    q = unsigned(floorf(x));
  return q & 0xFFu;
}

constexpr unsigned kHistBlocks = 256;

__global__ __launch_bounds__(256) void k_histogram(const float* __restrict__ v,
                                                   const float* __restrict__ mask, unsigned n,
                                                   float v_min, float step_inverse,
                                                   unsigned* __restrict__ partial) {
  __shared__ unsigned h[4][256];
  for (unsigned b = threadIdx.x; b < 4u * 256u; b += 256u) h[b / 256u][b % 256u] = 0u;
  __syncthreads();
  const unsigned wave = threadIdx.x >> 6;
  for (unsigned i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
    if (mask && mask[i] != 1.0f) continue;
    atomicAdd(&h[wave][quantize_bin(v[i], v_min, step_inverse)], 1u);
  }
  __syncthreads();
  const unsigned bin = threadIdx.x;
  partial[blockIdx.x * 256u + bin] = h[0][bin] + h[1][bin] + h[2][bin] + h[3][bin];
}

__global__ __launch_bounds__(256) void k_histogram_merge(const unsigned* __restrict__ partial,
                                                         unsigned blocks,
                                                         unsigned* __restrict__ hist) {
  const unsigned bin = threadIdx.x;
  unsigned s = 0;
  for (unsigned g = 0; g < blocks; ++g) s += partial[g * 256u + bin];
  hist[bin] = s;
}

// sum over bins of KLDivergence (runner.hpp:746-762): p_b / n_p
// log(p_b / max(q_b, q_lower_bound) * n_q / n_p), terms in FP32 as in the
// reference, summed in FP64 in bin order
__global__ __launch_bounds__(256) void k_kl_divergence(const unsigned* __restrict__ hp,
                                                       const unsigned* __restrict__ hq,
                                                       unsigned n_p, unsigned n_q,
                                                       float q_lower_bound, unsigned bins,
                                                       float* __restrict__ out) {
  __shared__ double part[256];
  const float n_p_inv = 1.0f / float(n_p);
  const float n_q_div_n_p = float(n_q) / float(n_p);
  double s = 0.0;
  for (unsigned b = threadIdx.x; b < bins; b += 256u) {
    const float fp = float(hp[b]), fq = float(hq[b]);
    if (fp != 0.0f) s += double(fp * n_p_inv * logf(fp / fmaxf(fq, q_lower_bound) * n_q_div_n_p));
  }
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (unsigned k = 0; k < 256u; ++k) t += part[k];
    out[0] = float(t);
  }
}

// 1 for every volume-map cell with a non-zero node
__global__ __launch_bounds__(256) void k_map_flags(GridGeomDev g, const float* __restrict__ nodes,
                                                   unsigned char* __restrict__ out) {
  const unsigned c = blockIdx.x * 256u + threadIdx.x;
  const unsigned ncell = g.res[0] * g.res[1] * g.res[2];
  if (c >= ncell) return;
  const int ip[3] = {int(c % g.res[0]), int((c / g.res[0]) % g.res[1]), int(c / (g.res[0] * g.res[1]))};
  unsigned cell[32];
  grid_cell_nodes(g.res, ip, cell);
  unsigned char nz = 0;
  for (int k = 0; k < 32; ++k) nz |= nodes[cell[k]] != 0.0f ? 1 : 0;
  out[c] = nz;
}

}  // namespace

void launch_map_flags(const GridGeomDev& g, const float* nodes, unsigned char* out,
                      hipStream_t s) {
  const unsigned ncell = g.res[0] * g.res[1] * g.res[2];
  if (ncell == 0) return;
  hipLaunchKernelGGL(k_map_flags, dim3((ncell + 255u) / 256u), dim3(256), 0, s, g, nodes, out);
}

void launch_sample_neighbors(const SimParams& p, const DevArrays& A, const float* sample_x,
                             unsigned ns, float4* out, unsigned* counts, hipStream_t s) {
  if (ns == 0) return;
  const unsigned* pc = p.pellet_mode ? A.pcell_start : nullptr;
  if (p.wrap_mask)
    hipLaunchKernelGGL(k_sample_neighbors<true>, dim3(sample_grid(ns)), dim3(kSampleBlock), 0, s,
                       p, A.cell_start, A.x, A.sorted_id, pc, A.px, A.psorted_id, A.cap,
                       sample_x, ns, out, counts, A.anomaly);
  else
    hipLaunchKernelGGL(k_sample_neighbors<false>, dim3(sample_grid(ns)), dim3(kSampleBlock), 0, s,
                       p, A.cell_start, A.x, A.sorted_id, pc, A.px, A.psorted_id, A.cap,
                       sample_x, ns, out, counts, A.anomaly);
}
void launch_sample_bead_pellet(const SimParams& p, const DevArrays& A, const float* sample_x,
                               unsigned ns, unsigned offset, unsigned max_beads, float4* bead,
                               unsigned* bead_counts, float4* pel, unsigned* pel_counts,
                               hipStream_t s) {
  if (ns == 0) return;
  const unsigned* pc = p.pellet_mode ? A.pcell_start : nullptr;
  if (p.wrap_mask)
    hipLaunchKernelGGL(k_sample_bead_pellet<true>, dim3(sample_grid(ns)), dim3(kSampleBlock), 0,
                       s, p, A.cell_start, A.x, A.sorted_id, pc, A.px, A.psorted_id, A.cap,
                       sample_x, ns, offset, max_beads, bead, bead_counts, pel, pel_counts,
                       A.anomaly);
  else
    hipLaunchKernelGGL(k_sample_bead_pellet<false>, dim3(sample_grid(ns)), dim3(kSampleBlock), 0,
                       s, p, A.cell_start, A.x, A.sorted_id, pc, A.px, A.psorted_id, A.cap,
                       sample_x, ns, offset, max_beads, bead, bead_counts, pel, pel_counts,
                       A.anomaly);
}
void launch_sample_with_pellets(const SimParams& p, int what, const float4* nbr,
                                const unsigned* counts, const float4* pel,
                                const unsigned* pcounts, unsigned K, unsigned ns,
                                const float* density, const float* v, const float* sample_v,
                                float* out, hipStream_t s) {
  if (ns == 0) return;
  hipLaunchKernelGGL(k_sample_with_pellets, dim3(sample_grid(ns)), dim3(kSampleBlock), 0, s, p,
                     what, nbr, counts, pel, pcounts, K, ns, density, v, sample_v, out);
}
void launch_pellet_sums(const SimParams& p, int what, const float4* bead,
                        const unsigned* bead_counts, const float4* pel,
                        const unsigned* pel_counts, unsigned K, unsigned n, float* out,
                        hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_pellet_sums, dim3(sample_grid(n)), dim3(kSampleBlock), 0, s, p, what, bead,
                     bead_counts, pel, pel_counts, K, n, out);
}
void launch_sample_boundary(const SimParams& p, const BoundaryDev* bodies, unsigned nb,
                            const float* sample_x, unsigned ns, float4* rec, float4* ker,
                            hipStream_t s) {
  if (ns == 0 || nb == 0) return;
  hipLaunchKernelGGL(k_sample_boundary, dim3(sample_grid(ns)), dim3(kSampleBlock), 0, s, p,
                     bodies, nb, sample_x, ns, rec, ker);
}
void launch_sample_fluid(const SimParams& p, const float4* nbr, const unsigned* counts,
                         unsigned K, unsigned ns, const float* density, const float* q,
                         unsigned comps, float* out, hipStream_t s) {
  if (ns == 0) return;
  hipLaunchKernelGGL(k_sample_fluid, dim3(sample_grid(ns)), dim3(kSampleBlock), 0, s, p, nbr,
                     counts, K, ns, density, q, comps, out);
}
void launch_sample_density(const SimParams& p, const float4* nbr, const unsigned* counts,
                           unsigned K, unsigned ns, const float4* ker, unsigned nb, float* out,
                           hipStream_t s) {
  if (ns == 0) return;
  hipLaunchKernelGGL(k_sample_density, dim3(sample_grid(ns)), dim3(kSampleBlock), 0, s, p, nbr,
                     counts, K, ns, ker, nb, out);
}
void launch_sample_velocity(const SimParams& p, const float4* nbr, const unsigned* counts,
                            unsigned K, unsigned ns, const float* density, const float* v,
                            const float* sample_v, const float4* rec, const float4* ker,
                            unsigned nb, const float* rx, const float* rv, const float* rw,
                            float* out, hipStream_t s) {
  if (ns == 0) return;
  if (sample_v)
    hipLaunchKernelGGL(k_sample_velocity<true>, dim3(sample_grid(ns)), dim3(kSampleBlock), 0, s, p,
                       nbr, counts, K, ns, density, v, sample_v, rec, ker, nb, rx, rv, rw, out);
  else
    hipLaunchKernelGGL(k_sample_velocity<false>, dim3(sample_grid(ns)), dim3(kSampleBlock), 0, s,
                       p, nbr, counts, K, ns, density, v, sample_v, rec, ker, nb, rx, rv, rw, out);
}
void launch_histogram256(const float* v, const float* mask, unsigned n, float v_min,
                         float step_inverse, unsigned* partial, unsigned* hist, hipStream_t s) {
  hipLaunchKernelGGL(k_histogram, dim3(kHistBlocks), dim3(256), 0, s, v, mask, n, v_min,
                     step_inverse, partial);
  hipLaunchKernelGGL(k_histogram_merge, dim3(1), dim3(256), 0, s, partial, kHistBlocks, hist);
}
unsigned histogram_partial_count() { return kHistBlocks * 256u; }

// compute_boundary_mask (runner.hpp:4368-4397): mask[i] is multiplied by
// (sign * d(R(q)^T (x_i - x_b)) >= threshold), d the boundary's signed
// distance (for a distance map the interpolated map, FLT_MAX outside it).
__global__ __launch_bounds__(256) void k_boundary_mask(BoundaryDev B, const float* __restrict__ x,
                                                      float threshold, float* __restrict__ mask,
                                                      unsigned n) {
  const unsigned i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const float4 q = make_float4(B.q[0], B.q[1], B.q[2], B.q[3]);
  const vec3 xi = mk(x[3 * size_t(i)], x[3 * size_t(i) + 1], x[3 * size_t(i) + 2]);
  const vec3 local = rotate(xi - mk(B.x[0], B.x[1], B.x[2]), conjugate(q));
  const float d = (B.type == ALV_DIST_MAP ? map_distance(B, local) : sdf_eval(B, local)) * B.sign;
  mask[i] *= (d >= threshold) ? 1.0f : 0.0f;
}
void launch_boundary_mask(const BoundaryDev& B, const float* x, float threshold, float* mask,
                          unsigned n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_boundary_mask, dim3((n + 255u) / 256u), dim3(256), 0, s, B, x, threshold,
                     mask, n);
}
void launch_kl_divergence(const unsigned* hp, const unsigned* hq, unsigned n_p, unsigned n_q,
                          float q_lower_bound, unsigned bins, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_kl_divergence, dim3(1), dim3(256), 0, s, hp, hq, n_p, n_q, q_lower_bound,
                     bins, out);
}

}  // namespace alv
