// alv_step.hip -- HIP kernels of the per-step SPH solver core for gfx950.
//
// Layout: particles are counting-sorted by grid cell every step (cell size h,
// z-fastest cell numbering as the reference's pid_length,
// variable.hpp:286-290; ascending particle id inside a cell).  All per-step
// passes run in that sorted order on SoA arrays, so a wavefront's 64 lanes
// touch neighbouring particles and the neighbour gathers stay in L2.  The
// neighbour list stores u32 sorted indices column-major ([slot][particle]) so
// that slot t of 64 lanes is one 256-byte coalesced load; displacements are
// recomputed from positions (bit-identical to the reference's stored x_ij,
// because positions do not change inside a step).
//
// Convergence: each residual pass writes one FP64 partial sum per block; a
// one-block finaliser reduces them in a fixed order and updates the device
// control block (iteration count, done flag) with the reference's
// min/max-iteration semantics (solver_df.hpp:131-133, 296-298).  Iteration
// kernels return at once when the flag is set, so the host can enqueue
// iterations in chunks without a round trip per iteration.

#include <float.h>

#include <algorithm>
#include <type_traits>

#include "alv_device.hpp"
#include "alv_kernels.hpp"

namespace alv {

#define ALV_TID() (xcd_block() * kBlock + threadIdx.x)

__device__ inline float std_max(float a, float b) { return (a < b) ? b : a; }

// Periodic wrap of a displacement, resolved at compile time (the reference's
// template <U wrap>, runner.hpp:1869) so that non-periodic neighbour loops
// carry no wrap branches.
template <bool W>
__device__ inline vec3 wrapd(vec3 d, const SimParams& p) {
  if constexpr (W) return wrap_disp(d, p);
  return d;
}

// Neighbour loops run in batches: the kNB indices, then all gathers of the
// batch, are issued before any arithmetic, so a wave waits for two memory
// latencies per batch instead of two per neighbour.  Empty slots point at the
// particle itself (a valid address whose values are ignored); contributions
// are still added in list order.
constexpr int kNB = 4;
constexpr unsigned kCB = 8;  // candidate batch of the neighbour search
__device__ inline void batch_ids(const DevArrays& A, unsigned s, unsigned nn, unsigned t0,
                                 unsigned (&j)[kNB]) {
  static_assert(kNB == 4, "one neighbour group per batch");
  uint4 g = A.nbr[nbr_group(A, t0 >> 2, s)];
  // The group stays one 16-byte load issued before the gathers: otherwise the
  // compiler may load each later slot on its own under its count test, with
  // a wait for each.
  asm volatile("" : "+v"(g.x), "+v"(g.y), "+v"(g.z), "+v"(g.w));
  j[0] = t0 + 0 < nn ? g.x : s;
  j[1] = t0 + 1 < nn ? g.y : s;
  j[2] = t0 + 2 < nn ? g.z : s;
  j[3] = t0 + 3 < nn ? g.w : s;
}

// slot t of a grouped list (nbr_group layout) of particle s
__device__ inline unsigned pellet_at(const uint4* list, const DevArrays& A, unsigned s,
                                     unsigned t) {
  return reinterpret_cast<const unsigned*>(list + nbr_group(A, t >> 2, s))[t & 3u];
}

// The w lane of the sorted position carries the stiffness of the next
// stiffness pass (kappa_v, kfac or kappa, written by the pass that computes
// it), so a stiffness pass gathers x_j and k_j with one load.
__device__ inline void set_kw(const DevArrays& A, unsigned s, float k) {
  reinterpret_cast<float*>(A.x + s)[3] = k;
}

// The w lane of the padded velocity carries the density, so the viscosity
// pass gathers v_j and rho_j with one load; velocity updates keep it.
__device__ inline void add_v(DevArrays& A, unsigned s, vec3 dv) {
  float4 v = A.v[s];
  A.v[s] = make_float4(v.x + dv.x, v.y + dv.y, v.z + dv.z, v.w);
}
__device__ inline float std_min(float a, float b) { return (b < a) ? b : a; }
// The viscosity's coefficient of a neighbour, mu m / rho_j, which
// compute_viscosity (runner.hpp:2150-2249) divides out for every pair: it is
// stored with the particle's velocity (v.w) when its density is computed, so
// one correctly rounded division per particle gives the quotient of every
// pair.
__device__ inline float visc_coefficient(const SimParams& p, float rho) {
  return (p.visc * p.mass) / rho;
}

// ---------------------------------------------------------------------------
// wave / block reductions (wave64)
// ---------------------------------------------------------------------------
__device__ inline double block_sum_d(double v, double* lds) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const unsigned lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  if (lane == 0) lds[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (unsigned i = 0; i < blockDim.x / 64u; ++i) r += lds[i];
  return r;
}
__device__ inline float block_max_f(float v, float* lds) {
  for (int o = 32; o > 0; o >>= 1) v = std_max(v, __shfl_down(v, o, 64));
  const unsigned lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  if (lane == 0) lds[w] = v;
  __syncthreads();
  float r = -FLT_MAX;
  if (threadIdx.x == 0)
    for (unsigned i = 0; i < blockDim.x / 64u; ++i) r = std_max(r, lds[i]);
  return r;
}
// Block size of the one-block-per-scene reductions over the per-block
// partials (finalisers, time step): four partials per thread at 1 M
// particles, so the loads of the block are in flight together.
constexpr unsigned kFinBlock = 1024;
// Publishes a control block that this thread has just updated to the host's
// copy (pinned, coherent, mapped): the host reads it after waiting for the
// stream, with no copy command.  The system-scope fence orders the stores
// before the kernel's completion.
__device__ inline void publish_ctl(const DevArrays& A, const StepCtl* ctl, unsigned scene) {
  A.ctl_pub[scene] = *ctl;
  __threadfence_system();
}
// exclusive scan of one u32 per thread over a block of 256 threads
__device__ inline unsigned block_excl_scan_u32(unsigned v, unsigned* lds, unsigned* total) {
  const unsigned lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  unsigned inc = v;
  for (int o = 1; o < 64; o <<= 1) {
    unsigned t = __shfl_up(inc, o, 64);
    if (lane >= unsigned(o)) inc += t;
  }
  if (lane == 63) lds[w] = inc;
  __syncthreads();
  unsigned off = 0, tot = 0;
  for (unsigned i = 0; i < blockDim.x / 64u; ++i) {
    unsigned t = lds[i];
    if (i < w) off += t;
    tot += t;
  }
  __syncthreads();
  if (total) *total = tot;
  return off + inc - v;
}

// ---------------------------------------------------------------------------
// Neighbour search: counting sort of particles by cell
// (replaces update_particle_grid, runner.hpp:1840-1867)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_cell_count(SceneTab T, DevArrays A) {
  ALV_SCENE_LINEAR(T, A, i);
  if (i >= p.first + p.n) return;
  vec3 x = ld3(A.ux, i);
  int ip[3];
  cell_ipos(x, p, ip);
  unsigned k;
  if (within_grid(ip, p)) {
    k = p.cell_base + cell_linear(ip, p);
  } else {
    k = p.cell_base + p.num_cells;  // outside the grid: own bin after the last cell
    A.anomaly[0] = 1u;
  }
  A.key[i] = k;
  A.slot[i] = atomicAdd(&A.cell_count[k], 1u);
}

constexpr unsigned kScanPerThread = kScanPerBlock / kBlock;
static_assert(kScanPerThread * kBlock == kScanPerBlock, "whole cells per thread");

__global__ __launch_bounds__(256) void k_scan_reduce(unsigned count, const unsigned* in,
                                                     unsigned* block_sums) {
  __shared__ double lds[4];
  const unsigned base = blockIdx.x * kScanPerBlock;
  unsigned s = 0;
  for (unsigned k = 0; k < kScanPerThread; ++k) {
    unsigned idx = base + k * kBlock + threadIdx.x;
    if (idx < count) s += in[idx];
  }
  double t = block_sum_d(double(s), lds);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = unsigned(t);
}

// single block: exclusive scan of the block sums, in place
__global__ __launch_bounds__(256) void k_scan_top(unsigned nblocks, unsigned* sums) {
  __shared__ unsigned lds[4];
  const unsigned per = (nblocks + kBlock - 1) / kBlock;
  const unsigned base = threadIdx.x * per;
  unsigned local = 0;
  for (unsigned k = 0; k < per; ++k)
    if (base + k < nblocks) local += sums[base + k];
  unsigned off = block_excl_scan_u32(local, lds, nullptr);
  for (unsigned k = 0; k < per; ++k) {
    if (base + k < nblocks) {
      unsigned v = sums[base + k];
      sums[base + k] = off;
      off += v;
    }
  }
}

// out[0..count] = exclusive scan of in[0..count); out[count] = total
__global__ __launch_bounds__(256) void k_scan_apply(unsigned count, const unsigned* in,
                                                    unsigned* out, const unsigned* offsets) {
  __shared__ unsigned lds[4];
  const unsigned base = blockIdx.x * kScanPerBlock + threadIdx.x * kScanPerThread;
  unsigned vals[kScanPerThread];
  unsigned local = 0;
#pragma unroll
  for (unsigned k = 0; k < kScanPerThread; ++k) {
    vals[k] = (base + k < count) ? in[base + k] : 0u;
    local += vals[k];
  }
  unsigned run = block_excl_scan_u32(local, lds, nullptr) + offsets[blockIdx.x];
#pragma unroll
  for (unsigned k = 0; k < kScanPerThread; ++k) {
    if (base + k < count) out[base + k] = run;
    run += vals[k];
    if (base + k + 1 == count) out[count] = run;
  }
}

__global__ __launch_bounds__(256) void k_scatter(SceneTab T, DevArrays A) {
  ALV_SCENE_LINEAR(T, A, i);
  if (i >= p.first + p.n) return;
  A.sort_tmp[A.cell_start[A.key[i]] + A.slot[i]] = i;
}

// Makes the order inside each cell canonical (ascending particle id): the
// atomic slots of k_cell_count are arrival-ordered, so every particle takes
// the position of its rank among its cell's particles in the scatter's
// copy.  One thread per sorted slot t: the scatter's copy holds particle
// sort_tmp[t] there, and its cell's slots [b, e) lie around t, so the rank
// loop and the write stay within a few cache lines of the thread's
// neighbours' (a thread per particle, in particle order, touches a random
// cell when particle order and space are unrelated).  The bin of particles
// outside the grid keeps the arrival order: nothing depends on it (those
// particles are in no cell and their own neighbour scans are independent).
__global__ __launch_bounds__(256) void k_cell_sort(SceneTab T, DevArrays A) {
  ALV_SCENE_LINEAR(T, A, t);
  if (t >= p.first + p.n) return;
  const unsigned i = A.sort_tmp[t];
  const unsigned k = A.key[i];
  if (k == p.cell_base + p.num_cells) {
    A.sorted_id[t] = i;
    return;
  }
  const unsigned b = A.cell_start[k];
  const unsigned e = A.cell_start[k + 1];
  if (e - b > p.cell_cap) A.anomaly[1] = 1u;
  unsigned rank = 0;
  for (unsigned u = b; u < e; ++u) rank += A.sort_tmp[u] < i ? 1u : 0u;
  A.sorted_id[b + rank] = i;
}

// LDS tile of every particle block.  For a block whose in-grid particles all
// lie in x-slab x0, the neighbours in slab xs (x0-1 .. x0+1) lie in the cell
// range from (xs, y0-1, z0-1) to (xs, y1+1, z1+1) in the linear cell order,
// (y0, z0) and (y1, z1) being the first and last particle's cells, clipped
// to the grid (with the whole first row when the block starts in row 0 and
// spans several rows, since row 1's neighbours reach (0, 0)).  Blocks that
// span several slabs, periodic grids and tiles over kTileMax fall back to
// global gathers.
__global__ __launch_bounds__(256) void k_tiles(SceneTab T, DevArrays A, unsigned nblocks) {
  const unsigned b = blockIdx.x * kBlock + threadIdx.x;
  if (b >= nblocks) return;
  Tile tl;
  for (int w = 0; w < 3; ++w) {
    tl.lo[w] = ~0u;
    tl.len[w] = 0u;
  }
  tl.total = 0u;
  tl.ok = 0u;
  const SimParams p = T.par[T.blk_scene[b]];
  const unsigned s0 = b * kBlock;
  const unsigned tail = A.cell_start[p.cell_base + p.num_cells];  // first particle outside the grid
  const unsigned end = min(p.first + p.n, tail);
  // no tile for periodic grids, for blocks outside the scene, and for a
  // block that holds particles outside the grid (their in-grid neighbour
  // cells need not lie in the window of the block's in-grid cells)
  if (p.wrap_mask != 0 || s0 < p.first || s0 >= end || min(s0 + kBlock, p.first + p.n) > tail) {
    A.tiles[b] = tl;
    return;
  }
  const unsigned s1 = min(s0 + kBlock, end) - 1u;
  const unsigned gx = unsigned(p.res[0]), gy = unsigned(p.res[1]), gz = unsigned(p.res[2]);
  const unsigned c0 = A.key[A.sorted_id[s0]] - p.cell_base;
  const unsigned c1 = A.key[A.sorted_id[s1]] - p.cell_base;
  const unsigned x0 = c0 / (gy * gz), x1 = c1 / (gy * gz);
  if (x0 != x1) {
    A.tiles[b] = tl;
    return;
  }
  const unsigned y0 = (c0 / gz) % gy, z0 = c0 % gz, y1 = (c1 / gz) % gy, z1 = c1 % gz;
  const unsigned ys = y0 > 0u ? y0 - 1u : 0u;
  const unsigned zs = (y0 > 0u || y1 == y0) ? (z0 > 0u ? z0 - 1u : 0u) : 0u;
  unsigned ye, ze;
  if (y1 + 1u < gy) {
    ye = y1 + 1u;
    ze = min(z1 + 1u, gz - 1u);
  } else {
    ye = gy - 1u;
    ze = gz - 1u;
  }
  unsigned k = 0, total = 0;
  for (int dx = -1; dx <= 1; ++dx) {
    const int xs = int(x0) + dx;
    if (xs < 0 || xs >= int(gx)) continue;
    const unsigned cs = (unsigned(xs) * gy + ys) * gz + zs;
    const unsigned ce = (unsigned(xs) * gy + ye) * gz + ze;
    const unsigned lo = A.cell_start[p.cell_base + cs];
    const unsigned hi = A.cell_start[p.cell_base + ce + 1u];
    tl.lo[k] = lo;
    tl.len[k] = hi - lo;
    total += hi - lo;
    ++k;
  }
  tl.total = total;
  tl.ok = total <= kTileMax ? 1u : 0u;
  A.tiles[b] = tl;
}

// ---------------------------------------------------------------------------
// Gather user -> sorted order and compute the boundary records
// (compute_particle_boundary_analytic<0>, runner.hpp:1982-2002, with
// compute_volume_and_boundary_x_analytic, runner.hpp:1738-1771)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gather_boundary(SceneTab T, DevArrays A, int kind) {
  ALV_SCENE(T, A, s);
  if (s >= p.first + p.n) return;
  const unsigned i = A.sorted_id[s];
  vec3 x = ld3(A.ux, i);
  vec3 v = ld3(A.uv, i);
  A.x[s] = f4(x, 0.0f);
  A.v[s] = f4(v, 0.0f);
  if (kind == ALV_KIND_DF || kind == ALV_KIND_DF_FACTOR) {
    A.kappa[s] = A.ukappa[i];
    A.kappa_v[s] = A.ukappa_v[i];
    if (kind == ALV_KIND_DF_FACTOR) A.alpha[s] = A.ualpha[i];
  } else if (kind == ALV_KIND_II) {
    A.pressure[s] = A.upressure[i];
  }
  if (p.pellet_mode) {
    // no boundary records: the pellet passes add the body forces
    for (unsigned b = 0; b < p.num_boundaries; ++b) {
      if (!A.bnd[b].accum) continue;
      const size_t o = size_t(b) * A.cap + s;
      A.bforce[o] = make_float4(0.f, 0.f, 0.f, 0.f);
      A.btorque[o] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    A.bmask[s] = 0u;
    return;
  }
  unsigned mask = 0;
  for (unsigned b = 0; b < p.num_boundaries; ++b) {
    const BoundaryDev& B = A.bnd[b];
    float4 rec, ker;
    const float vol = boundary_record<true>(B, x, p, rec, ker);
    // Records are stored only where Psi != 0; every contribution of a
    // boundary with Psi = 0 is an exact zero, so the passes skip it.
    if (vol != 0.0f) {
      const size_t o = size_t(b) * A.cap + s;
      A.brec[o] = rec;
      A.bker[o] = ker;
      mask |= 1u << b;
      if (B.accum) {
        A.bforce[o] = make_float4(0.f, 0.f, 0.f, 0.f);
        A.btorque[o] = make_float4(0.f, 0.f, 0.f, 0.f);
      }
    }
  }
  A.bmask[s] = mask;
}

// velocity of a rigid boundary at world point bx
__device__ inline vec3 rigid_velocity(const BoundaryDev& B, vec3 bx) {
  return cross(mk(B.omega[0], B.omega[1], B.omega[2]), bx - mk(B.x[0], B.x[1], B.x[2])) +
         mk(B.v[0], B.v[1], B.v[2]);
}

// Where the neighbour loops read x_j and v_j: global memory (L2 gathers),
// or the block's LDS tile.
struct GlobalXV {
  const float4* x;
  const float4* v;
  __device__ float4 X(unsigned j) const { return x[j]; }
  __device__ float4 V(unsigned j) const { return v[j]; }
};
// LDS slot of sorted particle j: the tile's ranges are in increasing order.
struct TileMap {
  unsigned lo1, lo2, base0, base1, base2;
  __device__ unsigned slot(unsigned j) const {
    return j >= lo2 ? j - base2 : (j >= lo1 ? j - base1 : j - base0);
  }
};
struct TileXV : TileMap {
  const float4* x;  // LDS
  const float4* v;  // LDS
  __device__ float4 X(unsigned j) const { return x[slot(j)]; }
  __device__ float4 V(unsigned j) const { return v[slot(j)]; }
};
// positions only (the neighbour search; stiffness passes read k_j in x_j.w)
struct GlobalX {
  const float4* x;
  __device__ float4 X(unsigned j) const { return x[j]; }
};

// The block's tile ranges and their LDS bases; false when the block has no
// usable tile.
__device__ inline bool tile_map(const DevArrays& A, unsigned blk, Tile& tl, TileMap& m) {
  if (!A.use_tiles) return false;
  tl = A.tiles[blk];
  if (!tl.ok) return false;
  unsigned off = 0, base[3];
#pragma unroll
  for (int w = 0; w < 3; ++w) {
    base[w] = tl.lo[w] - off;  // slot = j - base
    off += tl.len[w];
  }
  m.lo1 = tl.lo[1];
  m.lo2 = tl.lo[2];
  m.base0 = base[0];
  m.base1 = base[1];
  m.base2 = base[2];
  return true;
}

// Copies the block's tile of x and v into LDS (coalesced) when the block has
// one; every thread of the block must call it (it holds a barrier).  Returns
// false when the block uses global gathers instead.
__device__ inline bool stage_xv(const DevArrays& A, unsigned blk, TileXV& t) {
  __shared__ float4 lds_xv[2 * kTileMax];
  Tile tl;
  if (!tile_map(A, blk, tl, t)) return false;
  unsigned off = 0;
#pragma unroll
  for (int w = 0; w < 3; ++w) {
    const unsigned lo = tl.lo[w], len = tl.len[w];
    for (unsigned k = threadIdx.x; k < len; k += kBlock) {
      lds_xv[off + k] = A.x[lo + k];
      lds_xv[kTileMax + off + k] = A.v[lo + k];
    }
    off += len;
  }
  __syncthreads();
  t.x = lds_xv;
  t.v = lds_xv + kTileMax;
  return true;
}

// positions only (the neighbour search)
struct TileX : TileMap {
  const float4* x;  // LDS
  __device__ float4 X(unsigned j) const { return x[slot(j)]; }
};
__device__ inline bool stage_x(const DevArrays& A, unsigned blk, TileX& t) {
  __shared__ float4 lds_xo[kTileMax];
  Tile tl;
  if (!tile_map(A, blk, tl, t)) return false;
  unsigned off = 0;
#pragma unroll
  for (int w = 0; w < 3; ++w) {
    const unsigned lo = tl.lo[w], len = tl.len[w];
    for (unsigned k = threadIdx.x; k < len; k += kBlock) lds_xo[off + k] = A.x[lo + k];
    off += len;
  }
  __syncthreads();
  t.x = lds_xo;
  return true;
}

// ---------------------------------------------------------------------------
// Neighbour list (make_neighbor_list<wrap>, runner.hpp:1869-1911) fused with
// compute_density (runner.hpp:2052-2072) and, for DFSPH (mode 1),
// compute_dfsph_factor (runner.hpp:3124-3154); mode 2 also runs the
// divergence solve's warm start 0 (compute_density_change and
// warm_start_divergence_solve_0, runner.hpp:3192-3250) on the same walk over
// the list, reusing each neighbour's kernel gradient.  Cells are visited in
// the reference order (dx outer, dz inner); inside a cell in ascending
// particle id, which is ascending sorted index.
// ---------------------------------------------------------------------------
// Candidates staged per wave: the 64 particles of a wave mostly lie in one
// (x, y) column of cells, so for each (dx, dy) row their candidate spans lie
// in one range of about 64 + 2 cells' particles, which the wave copies into
// LDS with coalesced loads; the lanes then read their own spans from LDS
// instead of gathering from L2.  A row whose range exceeds kStage (a wave
// across two columns) reads global memory.  The stage keeps kCB slots of
// slack past kStage, so that a batch may read up to kCB - 1 entries past its
// span's end (masked out of its hits) without clamping its addresses.
constexpr unsigned kStage = 120;
// List slots per lane that the staged search keeps in LDS ([slot][lane]) until
// the density and factor pass, which stores them to the list four at a time
// for the whole wave (one coalesced 1 KB store per group); slots from
// kListLds on are written to the list during the search.
constexpr unsigned kListLds = 32;
static_assert(kListLds % 4u == 0u, "LDS slots end on a group boundary");

// Makes a wave's LDS writes visible to the wave's other lanes (the stage is
// private to a wave, so no workgroup barrier is needed).
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The staged instance holds 40 KB of LDS per block (4 blocks, 4 waves per
// SIMD); the others are bounded to 64 VGPRs for 8 waves per SIMD.
template <bool W, bool TL>
__global__ __launch_bounds__(256, (!W && !TL) ? 4 : 8) void k_neighbors(SceneTab T, DevArrays A,
                                                                          int with_factor) {
  // LDS staging and the LDS list for the gather instance without the
  // periodic wrap (its rows are single spans); the others keep the
  // lane-private scan
  constexpr bool kWaveStage = !W && !TL;
  __shared__ float4 stage_[kWaveStage ? 4 : 1][kWaveStage ? kStage + kCB : 1];
  __shared__ unsigned list_[kWaveStage ? 4 : 1][kWaveStage ? kListLds : 1][64];
  ALV_SCENE(T, A, s);
  // candidate positions from the block's LDS tile when it has one
  TileX tile_;
  bool tiled_ = false;
  if constexpr (TL) tiled_ = stage_x(A, s / kBlock, tile_);
  const bool live = s < p.first + p.n;
  if (!kWaveStage && !live) return;
  const unsigned K = p.kmax;
  // this lane's LDS list slots: slot t at lst[t * 64]
  unsigned* lst = &list_[kWaveStage ? threadIdx.x >> 6 : 0][0][kWaveStage ? threadIdx.x & 63u : 0];
  const vec3 xi = live ? ld3(A.x[s]) : mk(0.f, 0.f, 0.f);
  int ip[3];
  cell_ipos(xi, p, ip);
  unsigned cnt = 0;
  uint4 grp = make_uint4(0u, 0u, 0u, 0u);  // the group of slots being filled
  float dens = p.vol * p.w0;
  float sgk = 0.0f;
  vec3 gpi = mk(0.f, 0.f, 0.f);
  // Distances from x_i to the faces of its own cell.  A neighbour cell whose
  // box is farther than h holds no neighbour, so it is not scanned; skipped
  // cells contribute nothing to the list, whose order is unchanged.  The
  // reach carries a margin over rounding: a particle's cell comes from the
  // rounded x/h, so it can lie outside its cell's box by a few units in the
  // last place of its coordinate, and the face coordinate and the differences
  // are rounded too.  Those errors scale with the coordinates' magnitude
  // (about 2^-23 |x| in all), so the margin is 2^-20 of the largest
  // coordinate (plus h), with a floor of 5e-5 h for coordinates near the
  // origin.
  float lo_d[3], hi_d[3];
  float mag = p.h;
  {
    const float xs[3] = {xi.x, xi.y, xi.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float c0 = float(ip[a] + p.off[a]) * p.h;
      lo_d[a] = xs[a] - c0;
      hi_d[a] = c0 + p.h - xs[a];
      mag = fmaxf(mag, fabsf(xs[a]) + p.h);
    }
  }
  const float reach = p.h * 1.00005f + mag * 9.5367431640625e-7f;
  const float reach2 = reach * reach;
  // The particle itself lies in row 4 (dx = dy = 0), unless a periodic axis
  // has fewer than 3 cells and another row wraps onto its own.
  const bool self_any_row =
      W && (((p.wrap_mask & 1) && p.res[0] < 3) || ((p.wrap_mask & 2) && p.res[1] < 3));
  // Appends neighbour j at slot min(cnt, K-1) (the last slot keeps the
  // latest neighbour): LDS slots below kListLds in the staged instance, else
  // the register group, stored when complete.
  auto append = [&](unsigned j) {
    const unsigned slot = min(cnt, K - 1);
    if (kWaveStage && slot < kListLds) {
      lst[slot * 64u] = j;
    } else {
      const unsigned c4 = slot & 3u;
      grp.x = c4 == 0 ? j : grp.x;
      grp.y = c4 == 1 ? j : grp.y;
      grp.z = c4 == 2 ? j : grp.z;
      grp.w = c4 == 3 ? j : grp.w;
      if (c4 == 3u && cnt == slot) A.nbr[nbr_group(A, slot >> 2, s)] = grp;
    }
    ++cnt;
  };
  // candidates [jb, je) in batches: all positions of a batch are loaded
  // before any is tested (one memory latency per batch); the tests are
  // branch-free and give a bit mask of hits, which are then appended in
  // ascending j.  X(j0, u) is candidate j0 + u; kClamp: X must not be read
  // past je - 1 (global memory), else reads past the span are harmless.
  auto scan_span = [&](unsigned jb, unsigned je, auto&& X, auto clamp_tag, bool own_row) {
    constexpr bool kClamp = decltype(clamp_tag)::value;
    for (unsigned j0 = jb; j0 < je; j0 += kCB) {
      float4 xc[kCB];
#pragma unroll
      for (int u = 0; u < kCB; ++u) xc[u] = kClamp ? X(min(j0 + u, je - 1u), 0u) : X(j0, u);
      unsigned hits = 0u;
#pragma unroll
      for (int u = 0; u < kCB; ++u) {
        const vec3 d = wrapd<W>(xi - ld3(xc[u]), p);
        hits |= unsigned(length_sqr(d) < p.h2) << u;
      }
      // the batch's slots past the span and, in the particle's own row, the
      // particle itself, once per batch rather than per candidate
      hits &= __builtin_amdgcn_ubfe(0xffu, 0u, min(je - j0, kCB));
      if (own_row && s - j0 < kCB) hits &= ~(1u << (s - j0));
      if (kWaveStage && cnt + kCB <= min(kListLds, K)) {
        // Every candidate of the batch is written, without a branch, to the
        // slot after the batch's earlier hits: a miss is overwritten by the
        // next hit (or lies past the new count), so the slots cnt ..
        // cnt + hits - 1 end up holding the hits in ascending j.  The batch
        // stays below the LDS slots' end and below the cap K - 1, where the
        // latest neighbour replaces the last slot.
#pragma unroll
        for (int u = 0; u < kCB; ++u)
          lst[(cnt + unsigned(__builtin_popcount(hits & ((1u << u) - 1u)))) * 64u] = j0 + u;
        cnt += unsigned(__builtin_popcount(hits));
      } else {
        while (hits) {
          const unsigned j = j0 + unsigned(__builtin_ctz(hits));
          hits &= hits - 1u;
          append(j);
        }
      }
    }
  };
  using Clamp = std::integral_constant<bool, true>;
  using NoClamp = std::integral_constant<bool, false>;
  // The 27 cells in the reference order (dx outer, dz inner): for each
  // (dx, dy) row the up-to-3 cells along z are contiguous in the sorted
  // order unless a periodic wrap splits them, so each row is 1-3 spans.
  for (int r = 0; r < 9; ++r) {
    const int ddx = r / 3 - 1, ddy = r % 3 - 1;
    const float ex = ddx < 0 ? lo_d[0] : (ddx > 0 ? hi_d[0] : 0.0f);
    const float ey = ddy < 0 ? lo_d[1] : (ddy > 0 ? hi_d[1] : 0.0f);
    const float base = ex * ex + ey * ey;
    const int dz_lo = base + lo_d[2] * lo_d[2] < reach2 ? -1 : 0;
    const int dz_hi = base + hi_d[2] * hi_d[2] < reach2 ? 1 : 0;
    int nx = ip[0] + ddx, ny = ip[1] + ddy;
    if (p.wrap_mask & 1) nx = nx >= p.res[0] ? nx - p.res[0] : (nx < 0 ? nx + p.res[0] : nx);
    if (p.wrap_mask & 2) ny = ny >= p.res[1] ? ny - p.res[1] : (ny < 0 ? ny + p.res[1] : ny);
    const bool row_ok = live && base < reach2 && nx >= 0 && nx < p.res[0] && ny >= 0 &&
                        ny < p.res[1];
    const unsigned row = (unsigned(nx) * unsigned(p.res[1]) + unsigned(ny)) * unsigned(p.res[2]);
    if constexpr (kWaveStage) {
      // one span: cells nz_lo .. nz_hi of the row, clipped to the grid
      const int z0 = max(ip[2] + dz_lo, 0), z1 = min(ip[2] + dz_hi, p.res[2] - 1);
      unsigned jb = 0u, je = 0u;
      if (row_ok && z0 <= z1) {
        jb = A.cell_start[p.cell_base + row + unsigned(z0)];
        je = A.cell_start[p.cell_base + row + unsigned(z1) + 1u];
      }
      unsigned lo = jb < je ? jb : ~0u, hi = jb < je ? je : 0u;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, unsigned(__shfl_xor(int(lo), o, 64)));
        hi = max(hi, unsigned(__shfl_xor(int(hi), o, 64)));
      }
      const bool fits = hi > lo && hi - lo <= kStage;
      float4* st = stage_[threadIdx.x >> 6];
      wave_lds_sync();  // the previous row's reads are done
      if (fits)
        for (unsigned k = threadIdx.x & 63u; k < hi - lo; k += 64u) st[k] = A.x[lo + k];
      wave_lds_sync();
      if (jb < je) {
        if (fits)
          scan_span(jb, je, [&](unsigned j0, unsigned u) { return st[j0 - lo + u]; }, NoClamp(),
                    r == 4 || self_any_row);
        else
          scan_span(jb, je, [&](unsigned j, unsigned) { return A.x[j]; }, Clamp(),
                    r == 4 || self_any_row);
      }
    } else {
      if (!row_ok) continue;
      unsigned span_b[3], span_e[3], nspan = 0;
      int prev = -2;
      for (int dz = dz_lo; dz <= dz_hi; ++dz) {
        int nz = ip[2] + dz;
        if (p.wrap_mask & 4) nz = nz >= p.res[2] ? nz - p.res[2] : (nz < 0 ? nz + p.res[2] : nz);
        if (nz < 0 || nz >= p.res[2]) continue;
        const unsigned c = p.cell_base + row + unsigned(nz);
        if (nz == prev + 1) {
          span_e[nspan - 1] = A.cell_start[c + 1];
        } else {
          span_b[nspan] = A.cell_start[c];
          span_e[nspan] = A.cell_start[c + 1];
          ++nspan;
        }
        prev = nz;
      }
      for (unsigned q = 0; q < nspan; ++q) {
        if (tiled_)
          scan_span(span_b[q], span_e[q], [&](unsigned j, unsigned) { return tile_.X(j); },
                    Clamp(), r == 4 || self_any_row);
        else
          scan_span(span_b[q], span_e[q], [&](unsigned j, unsigned) { return A.x[j]; },
                    Clamp(), r == 4 || self_any_row);
      }
    }
  }
  if (!live) return;
  // The group holding the last slot is incomplete, or was rewritten by
  // overflowing neighbours after it was stored (register groups only; the
  // LDS slots are stored below).
  const unsigned nn = min(cnt, K);
  if (nn > 0 && (!kWaveStage || nn - 1u >= kListLds) && ((nn & 3u) != 0u || cnt > K))
    A.nbr[nbr_group(A, (nn - 1u) >> 2, s)] = grp;
  if (cnt > K) A.anomaly[2] = 1u;
  // Density and factor sums over the finished list, in list order.  The
  // list holds the first K-1 neighbours in search order and the last one
  // found in slot K-1, the order in which the reference accumulates them.
  // Doing the kernel arithmetic here, not inside the candidate scan, keeps
  // every lane of a wave busy: in the scan some lane finds a neighbour at
  // almost every candidate, so the whole wave would execute the kernel
  // evaluation for every candidate.  The lanes walk their lists in step, so
  // the LDS slots of group q are stored for the whole wave at once.
  const bool ws = with_factor == 2;  // divergence warm start 0 on this walk
  const vec3 vi = ws ? ld3(A.v[s]) : mk(0.f, 0.f, 0.f);
  float div = 0.0f;  // sum_j V (v_i - v_j) . gradW_ij, in list order
  for (unsigned t0 = 0; t0 < nn; t0 += kNB) {
    unsigned jn[kNB];
    if (kWaveStage && t0 < kListLds) {
      const uint4 g = make_uint4(lst[t0 * 64u], lst[(t0 + 1u) * 64u], lst[(t0 + 2u) * 64u],
                                 lst[(t0 + 3u) * 64u]);
      A.nbr[nbr_group(A, t0 >> 2, s)] = g;
      jn[0] = t0 + 0 < nn ? g.x : s;
      jn[1] = t0 + 1 < nn ? g.y : s;
      jn[2] = t0 + 2 < nn ? g.z : s;
      jn[3] = t0 + 3 < nn ? g.w : s;
    } else {
      batch_ids(A, s, nn, t0, jn);
    }
    float4 xj[kNB], vj[kNB];
#pragma unroll
    for (int u = 0; u < kNB; ++u) xj[u] = tiled_ ? tile_.X(jn[u]) : A.x[jn[u]];
    if (ws) {
#pragma unroll
      for (int u = 0; u < kNB; ++u) vj[u] = A.v[jn[u]];
    }
    // An empty slot holds the particle itself (d = 0): its W term is
    // replaced by +0 and its gradient and divergence terms are signed zeros,
    // so every slot is added without a branch (sums that start at +0 never
    // become -0).
#pragma unroll
    for (int u = 0; u < kNB; ++u) {
      vec3 d = wrapd<W>(xi - ld3(xj[u]), p);
      const float w = W_d2(length_sqr(d), p);
      dens += p.vol * (t0 + u < nn ? w : 0.0f);
      if (with_factor) {
        const vec3 g = gradW_rn(d, p);
        vec3 gp = (-p.vol) * g;
        sgk += length_sqr(gp);
        gpi -= gp;
        if (ws) div += p.vol * dot(vi - ld3(vj[u]), g);
      }
    }
  }
  unsigned mask = A.bmask[s];
  while (mask) {
    const unsigned b = __builtin_ctz(mask);
    mask &= mask - 1u;
    const size_t o = size_t(b) * A.cap + s;
    float4 bk = A.bker[o];
    dens += bk.w;
    gpi += mk(bk.x, bk.y, bk.z);
    if (ws) div += dot(vi - rigid_velocity(A.bnd[b], ld3(A.brec[o])), mk(bk.x, bk.y, bk.z));
  }
  A.nnb[s] = nn;
  const float rho = dens * p.rho0;
  A.density[s] = rho;
  reinterpret_cast<float*>(A.v + s)[3] = visc_coefficient(p, rho);
  if (with_factor) {
    sgk += length_sqr(gpi);
    A.alpha[s] = sgk > p.eps ? -1.0f / sgk : 0.0f;
  }
  if (ws) {
    // density_change, then the warm start of the stored stiffness
    float da = std_max(div, 0.0f);
    if (nn < 20u) da = 0.0f;
    const float dt = A.ctl->dt;
    const float kv_prev = A.kappa_v[s];
    const float kv = (da > 0.0f) ? 0.5f * std_max(kv_prev, -0.5f) / dt : 0.0f;
    A.kappa_v[s] = kv;
    set_kw(A, s, kv);
  }
}

// ---------------------------------------------------------------------------
// DFSPH building blocks
// ---------------------------------------------------------------------------

// sum_j V (v_i - v_j) . gradW_ij + sum_b (v_i - v_b) . Psi gradW_b
template <bool W, class Src>
__device__ inline float velocity_divergence(const SimParams& p, const DevArrays& A,
                                            unsigned s, vec3 xi, vec3 vi, unsigned nn,
                                            const Src& src) {
  float acc = 0.0f;
  for (unsigned t0 = 0; t0 < nn; t0 += kNB) {
    unsigned j[kNB];
    batch_ids(A, s, nn, t0, j);
    float4 xj[kNB], vj[kNB];
#pragma unroll
    for (int u = 0; u < kNB; ++u) {
      xj[u] = src.X(j[u]);
      vj[u] = src.V(j[u]);
    }
#pragma unroll
    for (int u = 0; u < kNB; ++u) {
      if (t0 + u < nn) {
        vec3 d = wrapd<W>(xi - ld3(xj[u]), p);
        acc += p.vol * dot(vi - ld3(vj[u]), gradW(d, p));
      }
    }
  }
  unsigned mask = A.bmask[s];
  while (mask) {
    const unsigned b = __builtin_ctz(mask);
    mask &= mask - 1u;
    const size_t o = size_t(b) * A.cap + s;
    float4 r = A.brec[o], k = A.bker[o];
    vec3 velj = rigid_velocity(A.bnd[b], ld3(r));
    acc += dot(vi - velj, mk(k.x, k.y, k.z));
  }
  return acc;
}

// compute_density_change (runner.hpp:3192-3228)
template <bool W, class Src>
__device__ inline float density_change(const SimParams& p, const DevArrays& A, unsigned s,
                                       const Src& src) {
  const vec3 xi = ld3(A.x[s]), vi = ld3(A.v[s]);
  const unsigned nn = A.nnb[s];
  float da = velocity_divergence<W>(p, A, s, xi, vi, nn, src);
  da = std_max(da, 0.0f);
  if (nn < 20u) da = 0.0f;
  return da;
}

// ---------------------------------------------------------------------------
// Several lanes per particle.  The gathers of the neighbour passes are bound
// by the vector L1's tag lookups: a wave64 gather is looked up four lanes at
// a time, one lookup per distinct cache line of the four.  With one particle
// per lane, the four lanes hold four particles whose lists drift apart, so
// their j-th neighbours lie on two or three lines.  With kLPP lanes per
// particle, lane q of a particle takes list slots kLPP k + q: consecutive
// slots, mostly consecutive particles of one cell, on fewer lines.  Each lane
// evaluates its slots' terms; every lane of the particle then adds the terms
// in slot order (DPP broadcasts within the quad), so the sum is the
// sequential sum over the list.
// ---------------------------------------------------------------------------
// The value of lane Q of this lane's particle.  Every lane of the quad is
// active where it is called, so bound_ctrl never applies; it lets the
// compiler fold the move into the add (v_add_f32_dpp).
template <int Q>
__device__ inline float lane_of_particle(float v) {
  constexpr int ctl = kLPP == 4u ? Q * 0x55 : (Q | Q << 2 | (2 + Q) << 4 | (2 + Q) << 6);
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), ctl, 0xf, 0xf, true));
}
// acc + (the particle's lanes' t, in lane order)
__device__ inline float add_in_slot_order(float acc, float t) {
  acc += lane_of_particle<0>(t);
  acc += lane_of_particle<1>(t);
  if constexpr (kLPP == 4u) {
    acc += lane_of_particle<2>(t);
    acc += lane_of_particle<3>(t);
  }
  return acc;
}
// Slots t0 + q and t0 + kLPP + q of particle s's list, or s for a slot past
// the count; t0 is a multiple of 2 kLPP below the count.  A group past the
// list's last one is read from the last one (a valid address) and ignored.
// The loads are issued before either slot is used (see batch_ids).
__device__ inline void lane_slots(const DevArrays& A, unsigned s, unsigned q, unsigned t0,
                                  unsigned nn, unsigned& ja, unsigned& jb) {
  if constexpr (kLPP == 4u) {
    const unsigned* g = reinterpret_cast<const unsigned*>(A.nbr + nbr_group(A, t0 >> 2, s)) + q;
    const unsigned gb = min((t0 >> 2) + 1u, nbr_groups(A.kmax) - 1u) - (t0 >> 2);
    ja = g[0];
    jb = g[gb * 256u];  // 64 groups of 4 slots per group index
    asm volatile("" : "+v"(ja), "+v"(jb));
  } else {
    uint4 g = A.nbr[nbr_group(A, t0 >> 2, s)];
    asm volatile("" : "+v"(g.x), "+v"(g.y), "+v"(g.z), "+v"(g.w));
    ja = q ? g.y : g.x;
    jb = q ? g.w : g.z;
  }
  ja = t0 + q < nn ? ja : s;
  jb = t0 + kLPP + q < nn ? jb : s;
}

// velocity_divergence over the fluid list, kLPP lanes per particle; the
// boundary terms are added by every lane of the particle.  A slot past the
// count holds the particle itself, whose term is an exact zero (x_i - x_i =
// +0, so gradW is a signed zero and so is the term), so every slot's term is
// added without a test: a sum that starts at +0 never becomes -0 under
// round-to-nearest, and x + (+-0) = x for every other x.
// V: also the fluid part of the viscosity, sum_j mu m / rho_j (x_ij .
// gradW_ij) / (|x_ij|^2 + 0.01 h^2) (v_i - v_j), into *visc (every lane of
// the particle holds it): the same gathers (mu m / rho_j in v_j.w) and the
// same gradW_ij, term for term viscosity_sum's, added in list order.
template <bool W, bool V = false>
__device__ inline float velocity_divergence_q(const SimParams& p, const DevArrays& A,
                                              unsigned s, unsigned q, vec3 xi, vec3 vi,
                                              unsigned nn, vec3* visc = nullptr) {
  float acc = 0.0f;
  vec3 av = mk(0.f, 0.f, 0.f);
  const float e2 = 0.01f * p.h2;
  for (unsigned t0 = 0; t0 < nn; t0 += 2u * kLPP) {
    unsigned ja, jb;
    lane_slots(A, s, q, t0, nn, ja, jb);
    const float4 xa = A.x[ja], va = A.v[ja], xb = A.x[jb], vb = A.v[jb];
    const vec3 da = wrapd<W>(xi - ld3(xa), p), db = wrapd<W>(xi - ld3(xb), p);
    const vec3 ga = gradW_rn(da, p), gb = gradW_rn(db, p);
    const vec3 ua = vi - ld3(va), ub = vi - ld3(vb);
    acc = add_in_slot_order(acc, p.vol * dot(ua, ga));
    acc = add_in_slot_order(acc, p.vol * dot(ub, gb));
    if constexpr (V) {
      // an empty slot's term is an exact zero here too (x_ij = 0)
      const vec3 wa = div_rn(va.w * dot(da, ga), length_sqr(da) + e2) * ua;
      const vec3 wb = div_rn(vb.w * dot(db, gb), length_sqr(db) + e2) * ub;
      av.x = add_in_slot_order(av.x, wa.x);
      av.y = add_in_slot_order(av.y, wa.y);
      av.z = add_in_slot_order(av.z, wa.z);
      av.x = add_in_slot_order(av.x, wb.x);
      av.y = add_in_slot_order(av.y, wb.y);
      av.z = add_in_slot_order(av.z, wb.z);
    }
  }
  if constexpr (V) *visc = av;
  unsigned mask = A.bmask[s];
  while (mask) {
    const unsigned b = __builtin_ctz(mask);
    mask &= mask - 1u;
    const size_t o = size_t(b) * A.cap + s;
    float4 r = A.brec[o], k = A.bker[o];
    vec3 velj = rigid_velocity(A.bnd[b], ld3(r));
    acc += dot(vi - velj, mk(k.x, k.y, k.z));
  }
  return acc;
}
template <bool W>
__device__ inline float density_change_q(const SimParams& p, const DevArrays& A, unsigned s,
                                         unsigned q) {
  const vec3 xi = ld3(A.x[s]), vi = ld3(A.v[s]);
  const unsigned nn = A.nnb[s];
  float da = velocity_divergence_q<W>(p, A, s, q, xi, vi, nn);
  da = std_max(da, 0.0f);
  if (nn < 20u) da = 0.0f;
  return da;
}

// The density-change passes run with one lane per particle when they read
// neighbours from the block's LDS tile (TL), else with kLPP lanes per
// particle (global gathers); q is the lane within the particle's lanes.
template <bool W, bool TL>
__device__ inline float pass_density_change(const SimParams& p, const DevArrays& A,
                                            unsigned s, unsigned q, bool tiled,
                                            const TileXV& tile) {
  if constexpr (TL)
    return tiled ? density_change<W>(p, A, s, tile)
                 : density_change<W>(p, A, s, GlobalXV{A.x, A.v});
  else
    return density_change_q<W>(p, A, s, q);
}

// compute_density_adv (runner.hpp:3493-3523)
template <bool W, class Src>
__device__ inline float density_adv(const SimParams& p, const DevArrays& A, unsigned s,
                                    float dt, const Src& src) {
  const vec3 xi = ld3(A.x[s]), vi = ld3(A.v[s]);
  const unsigned nn = A.nnb[s];
  float delta = velocity_divergence<W>(p, A, s, xi, vi, nn, src);
  return std_max(A.density[s] / p.rho0 + dt * delta, 1.0f);
}
template <bool W, bool TL>
__device__ inline float pass_density_adv(const SimParams& p, const DevArrays& A, unsigned s,
                                         unsigned q, float dt, bool tiled, const TileXV& tile) {
  if constexpr (TL) {
    return tiled ? density_adv<W>(p, A, s, dt, tile)
                 : density_adv<W>(p, A, s, dt, GlobalXV{A.x, A.v});
  } else {
    const vec3 xi = ld3(A.x[s]), vi = ld3(A.v[s]);
    const unsigned nn = A.nnb[s];
    float delta = velocity_divergence_q<W>(p, A, s, q, xi, vi, nn);
    return std_max(A.density[s] / p.rho0 + dt * delta, 1.0f);
  }
}

// Velocity change of one stiffness pass (warm_start_divergence_solve_1,
// divergence_solve_iteration, warm_start_pressure_solve1,
// pressure_solve_iteration: runner.hpp:3252-3295, 3316-3363, 3452-3491,
// 3543-3588).  kj[] holds every particle's stiffness of this pass.
template <bool W, bool kDivergenceForm, class Src>
__device__ inline vec3 stiffness_dv(const SimParams& p, const DevArrays& A, unsigned s,
                                    float k_i, const Src& src, float dt) {
  const vec3 xi = ld3(A.x[s]);
  const unsigned nn = A.nnb[s];
  vec3 dv = mk(0.f, 0.f, 0.f);
  for (unsigned t0 = 0; t0 < nn; t0 += kNB) {
    unsigned j[kNB];
    batch_ids(A, s, nn, t0, j);
    // one gather per neighbour: the pass's stiffness k_j rides in x_j.w
    float4 xj[kNB];
#pragma unroll
    for (int u = 0; u < kNB; ++u) xj[u] = src.X(j[u]);
    // Every slot's term is computed and a skipped one (past the count, or
    // |k_i + k_j| <= eps, where the reference adds nothing) is replaced by
    // +0, which leaves the sum's bits unchanged; the gradient's division is
    // the correctly rounded div_rn (the same bits as the IEEE division).
#pragma unroll
    for (int u = 0; u < kNB; ++u) {
      const float k_sum = k_i + xj[u].w;
      const bool use = t0 + u < nn && fabsf(k_sum) > p.eps;
      const vec3 d = wrapd<W>(xi - ld3(xj[u]), p);
      vec3 term;
      if (kDivergenceForm)
        term = (dt * k_sum) * ((-p.vol) * gradW_rn(d, p));
      else
        term = ((dt * k_sum) * p.vol) * gradW_rn(d, p);
      term = use ? term : mk(0.f, 0.f, 0.f);
      if (kDivergenceForm)
        dv -= term;
      else
        dv += term;
    }
  }
  if (fabsf(k_i) > p.eps) {
    unsigned mask = A.bmask[s];
    while (mask) {
      const unsigned b = __builtin_ctz(mask);
      mask &= mask - 1u;
      const size_t o = size_t(b) * A.cap + s;
      float4 k = A.bker[o];
      vec3 acc = k_i * (-mk(k.x, k.y, k.z));
      dv -= dt * acc;
      const BoundaryDev& B = A.bnd[b];
      if (B.accum) {
        vec3 f = p.mass * acc;
        vec3 tq = cross(ld3(A.brec[o]) - mk(B.x[0], B.x[1], B.x[2]), f);
        float4 F = A.bforce[o], T = A.btorque[o];
        A.bforce[o] = f4(ld3(F) + f, 0.0f);
        A.btorque[o] = f4(ld3(T) + tq, 0.0f);
      }
    }
  }
  return dv;
}

// sum_j mu m / rho_j (x_ij . gradW_ij) / (|x_ij|^2 + 0.01 h^2) (v_i - v_j)
template <bool W, class Src>
__device__ inline vec3 viscosity_sum(const SimParams& p, const DevArrays& A, unsigned s,
                                     vec3 xi, vec3 vi, unsigned nn, const Src& src) {
  vec3 da = mk(0.f, 0.f, 0.f);
  for (unsigned t0 = 0; t0 < nn; t0 += kNB) {
    unsigned j[kNB];
    batch_ids(A, s, nn, t0, j);
    float4 xj[kNB], vj[kNB];  // vj.w = mu m / rho_j
#pragma unroll
    for (int u = 0; u < kNB; ++u) {
      xj[u] = src.X(j[u]);
      vj[u] = src.V(j[u]);
    }
    // An empty slot holds the particle itself: x_ij = 0 gives a zero
    // numerator and a zero velocity difference, so its term is an exact zero
    // and every slot is added without a test (da never becomes -0).  The
    // divisor |x_ij|^2 + 0.01 h^2 is normal: div_rn.
#pragma unroll
    for (int u = 0; u < kNB; ++u) {
      vec3 xixj = wrapd<W>(xi - ld3(xj[u]), p);
      da += div_rn(vj[u].w * dot(xixj, gradW_rn(xixj, p)),
                   length_sqr(xixj) + 0.01f * p.h2) *
            (vi - ld3(vj[u]));
    }
  }
  return da;
}

__device__ inline void orthogonal_vectors(vec3 n, vec3& t1, vec3& t2) {
  vec3 v = mk(1.f, 0.f, 0.f);
  if (fabsf(dot(v, n)) > 0.999f) {
    v.x = 0.f;
    v.y = 1.f;
  }
  t1 = cross(n, v);
  t2 = cross(n, t1);
  t1 = t1 * (1.0f / sqrtf(length_sqr(t1)));
  t2 = t2 * (1.0f / sqrtf(length_sqr(t2)));
}

// The boundary terms of compute_viscosity: four points around the closest
// boundary point of every boundary the particle sees, added to the running
// sum da in the reference's order; with Force, also each body's reaction
// (its per-particle force and torque accumulators) where the body
// accumulates forces.
template <bool Force>
__device__ inline vec3 boundary_viscosity(const SimParams& p, const DevArrays& A, unsigned s,
                                          vec3 xi, vec3 vi, vec3 da) {
  unsigned mask = A.bmask[s];
  while (mask) {
    const unsigned b = __builtin_ctz(mask);
    mask &= mask - 1u;
    const size_t o = size_t(b) * A.cap + s;
    const BoundaryDev& B = A.bnd[b];
    float4 rec = A.brec[o];
    vec3 bx = ld3(rec);
    vec3 r_x = mk(B.x[0], B.x[1], B.x[2]);
    vec3 normal = bx - xi;
    float nl2 = length_sqr(normal);
    if (nl2 > 0.0f) {
      normal = normal * (1.0f / sqrtf(nl2));
      vec3 t1, t2;
      orthogonal_vectors(normal, t1, t2);
      float dist = p.h - sqrtf(nl2);
      vec3 xs[4] = {bx - t1 * dist, bx + t1 * dist, bx - t2 * dist, bx + t2 * dist};
      float vol = 0.25f * rec.w;
      vec3 as[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        vec3 xix = xi - xs[q];
        vec3 gw = gradW(xix, p);
        vec3 vq = rigid_velocity(B, xs[q]);
        as[q] = (p.bvisc * vol * dot(xix, gw) / (length_sqr(xix) + 0.01f * p.h2)) * (vi - vq);
      }
      da += as[0] + as[1] + as[2] + as[3];
      if (Force && B.accum) {
        vec3 f[4], tq[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          f[q] = (-p.mass) * as[q];
          tq[q] = cross(xs[q] - r_x, f[q]);
        }
        float4 F = A.bforce[o], T = A.btorque[o];
        A.bforce[o] = f4(ld3(F) + (f[0] + f[1] + f[2] + f[3]), 0.f);
        A.btorque[o] = f4(ld3(T) + (tq[0] + tq[1] + tq[2] + tq[3]), 0.f);
      }
    }
  }
  return da;
}

// a = g + da with da the fluid sum plus the boundary terms (no reactions),
// stored; returns the particle's CFL term |v + a dt|^2.
__device__ inline float viscosity_finish(const SimParams& p, const DevArrays& A, unsigned s,
                                         vec3 xi, vec3 vi, vec3 da, float dt) {
  da = boundary_viscosity<false>(p, A, s, xi, vi, da);
  const vec3 a = mk(p.g[0], p.g[1], p.g[2]) + da;
  A.a[s] = f4(a, 0.f);
  return length_sqr(vi + a * dt);
}

// The density change as pass_density_change and, over the same walk of the
// list, the fluid part of the viscosity (viscosity_sum's terms in its order;
// from an LDS tile, one lane per particle, the two sums walk the list
// separately).  xi and vi are the particle's position and velocity.
template <bool W, bool TL>
__device__ inline float pass_density_change_visc(const SimParams& p, const DevArrays& A,
                                                 unsigned s, unsigned q, bool tiled,
                                                 const TileXV& tile, vec3& xi, vec3& vi,
                                                 vec3& visc) {
  xi = ld3(A.x[s]);
  vi = ld3(A.v[s]);
  const unsigned nn = A.nnb[s];
  float da;
  if constexpr (TL) {
    if (tiled) {
      da = velocity_divergence<W>(p, A, s, xi, vi, nn, tile);
      visc = viscosity_sum<W>(p, A, s, xi, vi, nn, tile);
    } else {
      const GlobalXV g{A.x, A.v};
      da = velocity_divergence<W>(p, A, s, xi, vi, nn, g);
      visc = viscosity_sum<W>(p, A, s, xi, vi, nn, g);
    }
  } else {
    da = velocity_divergence_q<W, true>(p, A, s, q, xi, vi, nn, &visc);
  }
  da = std_max(da, 0.0f);
  if (nn < 20u) da = 0.0f;
  return da;
}

// warm_start_divergence_solve_0 (runner.hpp:3230-3250)
// F: also record the warm-start features of the particle
template <bool W, bool F, bool TL>
__global__ __launch_bounds__(TL ? kBlock : kLPP * kBlock) void k_ws_div0(SceneTab T, DevArrays A) {
  ALV_SCENE_PASS(T, A, s, q, TL);
  TileXV tile_;
  bool tiled_ = false;
  if constexpr (TL) tiled_ = stage_xv(A, s / kBlock, tile_);
  if (s >= p.first + p.n) return;
  const float dt = A.ctl->dt;
  const float dadv = pass_density_change<W, TL>(p, A, s, q, tiled_, tile_);
  if (q != 0u) return;
  const float kv_prev = A.kappa_v[s];
  const float kv = (dadv > 0.0f) ? 0.5f * std_max(kv_prev, -0.5f) / dt : 0.0f;
  A.kappa_v[s] = kv;
  set_kw(A, s, kv);
  if (F) {
    float psi = 0.0f;
    unsigned mask = A.bmask[s];
    while (mask) {
      const unsigned b = __builtin_ctz(mask);
      mask &= mask - 1u;
      psi += A.brec[size_t(b) * A.cap + s].w;
    }
    const vec3 v = ld3(A.v[s]);
    A.feat[2 * size_t(s)] =
        make_float4(A.density[s] / p.rho0 - 1.0f, dadv, float(A.nnb[s]), psi);
    A.feat[2 * size_t(s) + 1] = make_float4(kv_prev, A.kappa[s], sqrtf(length_sqr(v)), dt);
  }
}

// warm_start_divergence_solve_0 (runner.hpp:3230-3250) with the learned
// warm start fused in: the density change and the eight features are
// computed as in k_ws_div0, then the whole wave runs the network on its 64
// particles (matrix cores; the forward weight operands are staged in LDS once
// per block and shared by its four waves).  The network's two outputs are
// corrections to the reference's warm starts 0.5 max(kappa_v, -0.5) and
// 0.5 max(kappa, -2.5e-4) (stored form, from the previous step's stored
// stiffnesses); the corrected divergence stiffness is used here and the
// density one is kept for warm_start_pressure_solve0.  Converged
// stiffnesses are never positive (k = rho_dot alpha and (rho* - 1) alpha
// with rho_dot >= 0, rho* >= 1, alpha <= 0), so the results are clamped to
// [bound, 0] with the bounds of the reference's own warm-start clamps; a
// non-finite prediction falls back to the reference's warm start.  F: also
// log the features for training.
template <bool W, bool F>
__global__ __launch_bounds__(256) void k_ws_div0_learned(SceneTab T, DevArrays A, MlpDev M,
                                                         int use_divergence) {
  constexpr int kOps = kOpsT4;  // forward operands F1 .. F4
  __shared__ bf16x8 ops_lds[kOps];
  {
    const uint4* src = reinterpret_cast<const uint4*>(M.ops);
    uint4* dst = reinterpret_cast<uint4*>(ops_lds);
    for (int k = threadIdx.x; k < kOps; k += kBlock) dst[k] = src[k];
  }
  __syncthreads();
  ALV_SCENE(T, A, s);
  const bool live = s < p.first + p.n;
  const float dt = A.ctl->dt;
  float f[kMlpIn] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  float dadv = 0.0f, kv = 0.0f;
  if (live) {
    dadv = density_change<W>(p, A, s, GlobalXV{A.x, A.v});
    const float kv_prev = A.kappa_v[s];
    kv = (dadv > 0.0f) ? 0.5f * std_max(kv_prev, -0.5f) / dt : 0.0f;
    float psi = 0.0f;
    unsigned mask = A.bmask[s];
    while (mask) {
      const unsigned b = __builtin_ctz(mask);
      mask &= mask - 1u;
      psi += A.brec[size_t(b) * A.cap + s].w;
    }
    const vec3 v = ld3(A.v[s]);
    f[0] = A.density[s] / p.rho0 - 1.0f;
    f[1] = dadv;
    f[2] = float(A.nnb[s]);
    f[3] = psi;
    f[4] = kv_prev;
    f[5] = A.kappa[s];
    f[6] = sqrtf(length_sqr(v));
    f[7] = dt;
    if (F) {
      A.feat[2 * size_t(s)] = make_float4(f[0], f[1], f[2], f[3]);
      A.feat[2 * size_t(s) + 1] = make_float4(f[4], f[5], f[6], f[7]);
    }
  }
  MlpDev Ml = M;
  Ml.ops = ops_lds;
  float y0, y1;
  mlp_infer_wave(Ml, f, y0, y1);  // every lane of the wave, live or not
  if (!live) return;
  // The network predicts corrections to the reference's warm starts (in
  // stored form): an uncertain network (outputs near 0) falls back to them.
  if (use_divergence && dadv > 0.0f && isfinite(y0))
    kv = std_min(std_max(0.5f * std_max(f[4], -0.5f) + y0, -0.5f), 0.0f) / dt;
  A.kappa_v[s] = kv;
  set_kw(A, s, kv);
  A.kpred[s] =
      isfinite(y1) ? std_min(std_max(0.5f * std_max(f[5], -0.00025f) + y1, -0.00025f), 0.0f) : y1;
}

// warm_start_divergence_solve_1 (runner.hpp:3252-3295)
template <bool W, bool TL>
__global__ __launch_bounds__(256) void k_ws_div1(SceneTab T, DevArrays A) {
  ALV_SCENE(T, A, s);
  TileX tk_;
  bool tiled_ = false;
  if constexpr (TL) tiled_ = stage_x(A, s / kBlock, tk_);
  if (s >= p.first + p.n) return;
  const float dt = A.ctl->dt;
  vec3 dv = tiled_ ? stiffness_dv<W, true>(p, A, s, A.kappa_v[s], tk_, dt)
                   : stiffness_dv<W, true>(p, A, s, A.kappa_v[s], GlobalX{A.x}, dt);
  add_v(A, s, dv);
}
// compute_velocity_of_density_change (runner.hpp:3297-3314); also stores the
// first iteration's stiffness k = rho_adv * alpha.  The density-change passes
// of the divergence solve also compute the viscosity (compute_viscosity's
// a = g + fluid and boundary terms, without the bodies' reactions) and the
// per-block CFL maxima from the velocities they see, wherever the pass may be
// the solve's last: the last pass that runs sees the final velocities, which
// k_viscosity would read.  This pass is the last only if the control block
// says no iteration follows; an iteration's k_div_err is not the last while
// the iteration count stays below the minimum and the maximum.  Both passes
// are bounded to 64 VGPRs (8 waves per SIMD); unbounded, the viscosity terms
// take 66 (7 waves per SIMD), which measured 2 % slower.
template <bool W, bool TL>
__global__ __launch_bounds__(TL ? kBlock : kLPP * kBlock, 8) void k_vel_density_change(SceneTab T, DevArrays A) {
  __shared__ float lds[kLPP * kBlock / 64];
  ALV_SCENE_PASS(T, A, s, q, TL);
  TileXV tile_;
  bool tiled_ = false;
  if constexpr (TL) tiled_ = stage_xv(A, s / kBlock, tile_);
  const float dt = A.ctl->dt;
  const bool last = A.ctl->div_done != 0u;  // the same for the whole block
  float cfl = -FLT_MAX;
  if (s < p.first + p.n) {
    vec3 xi, vi, visc;
    const float dadv =
        last ? pass_density_change_visc<W, TL>(p, A, s, q, tiled_, tile_, xi, vi, visc)
             : pass_density_change<W, TL>(p, A, s, q, tiled_, tile_);
    if (q == 0u) {
      A.rho_adv[s] = dadv;
      float inv_dt = 1.0f / dt;
      float alpha = A.alpha[s] * inv_dt;
      A.alpha[s] = alpha;
      const float k = dadv * alpha;
      A.kfac[s] = k;
      set_kw(A, s, k);
      if (last) cfl = viscosity_finish(p, A, s, xi, vi, visc, dt);
    }
  }
  if (last) {
    const float m = block_max_f(cfl, lds);
    if (threadIdx.x == 0) A.fpartials[qblk_] = m;
  }
}
// divergence_solve_iteration (runner.hpp:3316-3363)
template <bool W, bool TL>
__global__ __launch_bounds__(256) void k_div_iter(SceneTab T, DevArrays A) {
  ALV_SCENE(T, A, s);
  if (A.ctl->div_done) return;
  TileX tk_;
  bool tiled_ = false;
  if constexpr (TL) tiled_ = stage_x(A, s / kBlock, tk_);
  if (s >= p.first + p.n) return;
  const float dt = A.ctl->dt;
  const float k_i = A.kfac[s];
  A.kappa_v[s] += k_i;
  vec3 dv = tiled_ ? stiffness_dv<W, true>(p, A, s, k_i, tk_, dt)
                   : stiffness_dv<W, true>(p, A, s, k_i, GlobalX{A.x}, dt);
  add_v(A, s, dv);
}
// compute_divergence_solve_density_error (runner.hpp:3396-3410) + FP64
// partial sums of the residual
// (and the viscosity, as k_vel_density_change)
template <bool W, bool TL>
__global__ __launch_bounds__(TL ? kBlock : kLPP * kBlock, 8) void k_div_err(SceneTab T, DevArrays A) {
  __shared__ double lds[kLPP * kBlock / 64];
  __shared__ float ldsm[kLPP * kBlock / 64];
  ALV_SCENE_PASS(T, A, s, q, TL);
  if (A.ctl->div_done) return;
  TileXV tile_;
  bool tiled_ = false;
  if constexpr (TL) tiled_ = stage_xv(A, s / kBlock, tile_);
  const float dt = A.ctl->dt;
  // k_div_finalize continues while the count it reaches is below both bounds
  const unsigned it = A.ctl->div_iter + 1u;
  const bool last = !(it < A.ctl->min_div && it < A.ctl->max_div);
  double part = 0.0;
  float cfl = -FLT_MAX;
  if (s < p.first + p.n) {
    vec3 xi, vi, visc;
    const float dadv =
        last ? pass_density_change_visc<W, TL>(p, A, s, q, tiled_, tile_, xi, vi, visc)
             : pass_density_change<W, TL>(p, A, s, q, tiled_, tile_);
    if (q == 0u) {
      A.rho_adv[s] = dadv;
      const float k = dadv * A.alpha[s];
      A.kfac[s] = k;
      set_kw(A, s, k);
      part = double(dadv);
      if (last) cfl = viscosity_finish(p, A, s, xi, vi, visc, dt);
    }
  }
  const double t = block_sum_d(part, lds);
  if (threadIdx.x == 0) A.partials[qblk_] = t;
  if (last) {
    const float m = block_max_f(cfl, ldsm);
    if (threadIdx.x == 0) A.fpartials[qblk_] = m;
  }
}
__global__ __launch_bounds__(kFinBlock) void k_div_finalize(SceneTab T, DevArrays A) {
  __shared__ double lds[kFinBlock / 64];
  ALV_SCENE_BLOCK(T, A);
  StepCtl* ctl = A.ctl;
  if (ctl->div_done) return;
  double part = 0.0;
  for (unsigned b = blk_begin + threadIdx.x; b < blk_end; b += blockDim.x) part += A.partials[b];
  double t = block_sum_d(part, lds);
  if (threadIdx.x == 0) {
    float mean = float(t / double(p.n));
    unsigned it = ctl->div_iter + 1u;
    ctl->div_iter = it;
    ctl->mean_density_change = mean;
    bool again = (mean > ctl->tol_div / ctl->dt || it < ctl->min_div) && it < ctl->max_div;
    ctl->div_done = again ? 0u : 1u;
    publish_ctl(A, ctl, scene_);
  }
}

// Control block reset before the solves of a step (k_clear_grid, the step's
// first kernel; the neighbour phase reads only dt from the block).
__device__ inline void reset_ctl(const DevArrays& A, unsigned scene) {
  StepCtl* ctl = A.ctl + scene;
  ctl->div_iter = 0u;
  ctl->dens_iter = 0u;
  const float mx = FLT_MAX;
  bool div_go = (mx > ctl->tol_div / ctl->dt || 0u < ctl->min_div) && 0u < ctl->max_div;
  ctl->div_done = div_go ? 0u : 1u;
  ctl->mean_density_change = mx;
  ctl->mean_density_error = mx;
  ctl->dens_done = 0u;  // decided after update_dt (tolerance does not use dt)
  bool dens_go = (mx > ctl->tol_dens || 0u < ctl->min_dens) && 0u < ctl->max_dens;
  ctl->dens_done = dens_go ? 0u : 1u;
  publish_ctl(A, ctl, scene);
}

// ---------------------------------------------------------------------------
// Viscosity (compute_viscosity, runner.hpp:2150-2249) with clear_acceleration
// (a = g, runner.hpp:1974-1980) and the per-block maxima of calculate_cfl_v2
// (runner.hpp:2533-2541).  With the divergence solve on, the DFSPH step
// computes it in that solve's density-change passes instead (the last one that
// runs sees the final velocities) and adds the bodies' reactions in
// k_integrate_np; k_viscosity serves the step without a divergence solve and
// IISPH.
// ---------------------------------------------------------------------------
template <bool W, bool TL>
__global__ __launch_bounds__(256) void k_viscosity(SceneTab T, DevArrays A) {
  __shared__ float lds[4];
  ALV_SCENE(T, A, s);
  TileXV tile_;
  bool tiled_ = false;
  if constexpr (TL) tiled_ = stage_xv(A, s / kBlock, tile_);
  const float dt = A.ctl->dt;
  float cfl = -FLT_MAX;
  if (s < p.first + p.n) {
    const vec3 xi = ld3(A.x[s]), vi = ld3(A.v[s]);
    const unsigned nn = A.nnb[s];
    vec3 da = tiled_ ? viscosity_sum<W>(p, A, s, xi, vi, nn, tile_)
                     : viscosity_sum<W>(p, A, s, xi, vi, nn, GlobalXV{A.x, A.v});
    da = boundary_viscosity<true>(p, A, s, xi, vi, da);
    vec3 a = mk(p.g[0], p.g[1], p.g[2]) + da;
    A.a[s] = f4(a, 0.f);
    cfl = length_sqr(vi + a * dt);
  }
  float m = block_max_f(cfl, lds);
  if (threadIdx.x == 0) A.fpartials[xcd_block()] = m;
}

// Solver::update_dt (solver.hpp:153-169) on the device.
__global__ __launch_bounds__(kFinBlock) void k_update_dt(SceneTab T, DevArrays A) {
  __shared__ float lds[kFinBlock / 64];
  ALV_SCENE_BLOCK(T, A);
  float m = -FLT_MAX;
  for (unsigned b = blk_begin + threadIdx.x; b < blk_end; b += blockDim.x) m = std_max(m, A.fpartials[b]);
  float pm = block_max_f(m, lds);
  if (threadIdx.x == 0) {
    StepCtl* ctl = A.ctl;
    float pile_max = -FLT_MAX;
    for (unsigned b = 0; b < p.num_boundaries; ++b) {
      const BoundaryDev& B = A.bnd[b];
      float v2 = length_sqr(cross(mk(B.omega[0], B.omega[1], B.omega[2]),
                                  mk(B.max_distance, 0.f, 0.f)) +
                            mk(B.v[0], B.v[1], B.v[2]));
      if (v2 > pile_max) pile_max = v2;
    }
    float max_v2 = std_max(pm, pile_max);
    float max_v = sqrtf(max_v2);
    float ls = ctl->length_scale;
    float cfl_dt = ctl->cfl * ls / max_v;
    float dt = std_max(std_min(cfl_dt, ctl->max_dt), ctl->min_dt);
    ctl->particle_max_v2 = pm;
    ctl->pile_max_v2 = pile_max;
    ctl->max_v2 = max_v2;
    ctl->cfl_dt = cfl_dt;
    ctl->div_dt = ctl->dt;
    ctl->dt = dt;
    ctl->utilized_cfl = dt * max_v / ls;
    publish_ctl(A, ctl, scene_);
  }
}

// integrate_non_pressure_acceleration (runner.hpp:3422-3428).  With fin (a
// DFSPH step whose divergence solve computed the viscosity), first the rest
// of what follows that solve in the reference: divergence_solve_finish
// (runner.hpp:3412-3420) with the time step the solve ran with, and the
// bodies' viscous reactions (compute_viscosity), from the same velocities,
// which this kernel has not yet changed.
__global__ __launch_bounds__(256) void k_integrate_np(SceneTab T, DevArrays A, int fin) {
  ALV_SCENE(T, A, s);
  if (s >= p.first + p.n) return;
  const float dt = A.ctl->dt;
  if (fin) {
    const float dts = A.ctl->div_dt;
    A.alpha[s] *= dts;
    A.kappa_v[s] *= dts;
    if (A.bmask[s])
      boundary_viscosity<true>(p, A, s, ld3(A.x[s]), ld3(A.v[s]), mk(0.f, 0.f, 0.f));
  }
  add_v(A, s, dt * ld3(A.a[s]));
}

// warm_start_pressure_solve0 (runner.hpp:3430-3450)
// L: learned warm start (prediction from k_ws_div0_learned)
template <bool W, bool L, bool TL>
__global__ __launch_bounds__(TL ? kBlock : kLPP * kBlock) void k_ws_p0(SceneTab T, DevArrays A) {
  ALV_SCENE_PASS(T, A, s, q, TL);
  TileXV tile_;
  bool tiled_ = false;
  if constexpr (TL) tiled_ = stage_xv(A, s / kBlock, tile_);
  if (s >= p.first + p.n) return;
  const float dt = A.ctl->dt;
  const float dadv = pass_density_adv<W, TL>(p, A, s, q, dt, tiled_, tile_);
  if (q != 0u) return;
  float k = (dadv > 1.0f) ? (0.5f * std_max(A.kappa[s], -0.00025f) / (dt * dt)) : 0.0f;
  if (L && dadv > 1.0f) {
    const float kl = A.kpred[s] / (dt * dt);
    if (isfinite(kl)) k = kl;
  }
  A.kappa[s] = k;
  set_kw(A, s, k);
}
// warm_start_pressure_solve1 (runner.hpp:3452-3491)
template <bool W, bool TL>
__global__ __launch_bounds__(256) void k_ws_p1(SceneTab T, DevArrays A) {
  ALV_SCENE(T, A, s);
  TileX tk_;
  bool tiled_ = false;
  if constexpr (TL) tiled_ = stage_x(A, s / kBlock, tk_);
  if (s >= p.first + p.n) return;
  const float dt = A.ctl->dt;
  vec3 dv = tiled_ ? stiffness_dv<W, false>(p, A, s, A.kappa[s], tk_, dt)
                   : stiffness_dv<W, false>(p, A, s, A.kappa[s], GlobalX{A.x}, dt);
  add_v(A, s, dv);
}
// compute_rho_adv (runner.hpp:3525-3541); stores k = (rho_adv - 1) * alpha
template <bool W, bool TL>
__global__ __launch_bounds__(TL ? kBlock : kLPP * kBlock) void k_rho_adv(SceneTab T, DevArrays A) {
  ALV_SCENE_PASS(T, A, s, q, TL);
  TileXV tile_;
  bool tiled_ = false;
  if constexpr (TL) tiled_ = stage_xv(A, s / kBlock, tile_);
  if (s >= p.first + p.n) return;
  const float dt = A.ctl->dt;
  const float dadv = pass_density_adv<W, TL>(p, A, s, q, dt, tiled_, tile_);
  if (q != 0u) return;
  A.rho_adv[s] = dadv;
  float alpha = A.alpha[s] * (1.0f / (dt * dt));
  A.alpha[s] = alpha;
  const float k = (dadv - 1.0f) * alpha;
  A.kfac[s] = k;
  set_kw(A, s, k);
}
// pressure_solve_iteration (runner.hpp:3543-3588)
template <bool W, bool TL>
__global__ __launch_bounds__(256) void k_p_iter(SceneTab T, DevArrays A) {
  ALV_SCENE(T, A, s);
  if (A.ctl->dens_done) return;
  TileX tk_;
  bool tiled_ = false;
  if constexpr (TL) tiled_ = stage_x(A, s / kBlock, tk_);
  if (s >= p.first + p.n) return;
  const float dt = A.ctl->dt;
  const float k_i = A.kfac[s];
  A.kappa[s] += k_i;
  vec3 dv = tiled_ ? stiffness_dv<W, false>(p, A, s, k_i, tk_, dt)
                   : stiffness_dv<W, false>(p, A, s, k_i, GlobalX{A.x}, dt);
  add_v(A, s, dv);
}
// compute_pressure_solve_density_error (runner.hpp:3590-3604) + FP64
// partial sums of (rho_adv - 1)
template <bool W, bool TL>
__global__ __launch_bounds__(TL ? kBlock : kLPP * kBlock) void k_p_err(SceneTab T, DevArrays A) {
  __shared__ double lds[kLPP * kBlock / 64];
  ALV_SCENE_PASS(T, A, s, q, TL);
  if (A.ctl->dens_done) return;
  TileXV tile_;
  bool tiled_ = false;
  if constexpr (TL) tiled_ = stage_xv(A, s / kBlock, tile_);
  const float dt = A.ctl->dt;
  double part = 0.0;
  if (s < p.first + p.n) {
    const float dadv = pass_density_adv<W, TL>(p, A, s, q, dt, tiled_, tile_);
    if (q == 0u) {
      A.rho_adv[s] = dadv;
      const float k = (dadv - 1.0f) * A.alpha[s];
      A.kfac[s] = k;
      set_kw(A, s, k);
      part = double(dadv - 1.0f);
    }
  }
  double t = block_sum_d(part, lds);
  if (threadIdx.x == 0) A.partials[qblk_] = t;
}
__global__ __launch_bounds__(kFinBlock) void k_p_finalize(SceneTab T, DevArrays A) {
  __shared__ double lds[kFinBlock / 64];
  ALV_SCENE_BLOCK(T, A);
  StepCtl* ctl = A.ctl;
  if (ctl->dens_done) return;
  double part = 0.0;
  for (unsigned b = blk_begin + threadIdx.x; b < blk_end; b += blockDim.x) part += A.partials[b];
  double t = block_sum_d(part, lds);
  if (threadIdx.x == 0) {
    float mean = float(t / double(p.n));
    unsigned it = ctl->dens_iter + 1u;
    ctl->dens_iter = it;
    ctl->mean_density_error = mean;
    bool again = (mean > ctl->tol_dens || it < ctl->min_dens) && it < ctl->max_dens;
    ctl->dens_done = again ? 0u : 1u;
    publish_ctl(A, ctl, scene_);
  }
}

// pressure_solve_finish + integrate_velocity<wrap> (runner.hpp:3606-3623),
// then the state goes back to user order.
template <bool W>
__global__ __launch_bounds__(256) void k_df_finish(SceneTab T, DevArrays A, int dens_finish) {
  ALV_SCENE(T, A, s);
  if (s >= p.first + p.n) return;
  const float dt = A.ctl->dt;
  const unsigned i = A.sorted_id[s];
  vec3 v = ld3(A.v[s]);
  vec3 x = wrapd<W>(ld3(A.x[s]) + dt * v, p);
  float kappa = A.kappa[s];
  if (dens_finish) kappa *= dt * dt;
  st3(A.ux, i, x);
  st3(A.uv, i, v);
  A.ukappa[i] = kappa;
  A.ukappa_v[i] = A.kappa_v[s];
}

// compute_velocity_of_density_change (runner.hpp:3297-3314), the density
// change sum alone (no stiffness factor update): rho_adv <- rho_dot.
template <bool W, bool TL>
__global__ __launch_bounds__(TL ? kBlock : kLPP * kBlock) void k_density_change(SceneTab T, DevArrays A) {
  ALV_SCENE_PASS(T, A, s, q, TL);
  TileXV tile_;
  bool tiled_ = false;
  if constexpr (TL) tiled_ = stage_xv(A, s / kBlock, tile_);
  if (s >= p.first + p.n) return;
  const float dadv = pass_density_change<W, TL>(p, A, s, q, tiled_, tile_);
  if (q == 0u) A.rho_adv[s] = dadv;
}

// Training pairs of the step just taken, in sorted order: the features
// recorded at its divergence warm start and the converged stored stiffness
// (kappa_v, kappa) it ended with.
__global__ __launch_bounds__(256) void k_export_pairs(SceneTab T, DevArrays A,
                                                      float* __restrict__ x,
                                                      float* __restrict__ y) {
  ALV_SCENE_LINEAR(T, A, s);
  if (s >= p.first + p.n) return;
  const float4 f0 = A.feat[2 * size_t(s)], f1 = A.feat[2 * size_t(s) + 1];
  float* xo = x + size_t(s) * kMlpIn;
  xo[0] = f0.x; xo[1] = f0.y; xo[2] = f0.z; xo[3] = f0.w;
  xo[4] = f1.x; xo[5] = f1.y; xo[6] = f1.z; xo[7] = f1.w;
  // targets: the converged stored stiffnesses minus the reference's warm
  // starts from the previous ones (features 4 and 5), which the learned warm
  // start adds back
  const unsigned i = A.sorted_id[s];
  y[size_t(s) * 2 + 0] = A.ukappa_v[i] - 0.5f * std_max(f1.x, -0.5f);
  y[size_t(s) * 2 + 1] = A.ukappa[i] - 0.5f * std_max(f1.y, -0.00025f);
}

// integrate_velocity<wrap> (runner.hpp:3613-3623) on the user-order state
// x <- x + dt v, wrapped on periodic axes.
template <bool W>
__global__ __launch_bounds__(256) void k_integrate_velocity(SceneTab T, DevArrays A, float dt) {
  ALV_SCENE_LINEAR(T, A, i);
  if (i >= p.first + p.n) return;
  vec3 x = wrapd<W>(ld3(A.ux, i) + dt * ld3(A.uv, i), p);
  st3(A.ux, i, x);
}

// ---------------------------------------------------------------------------
// IISPH (runner.hpp:2609-2887)
// ---------------------------------------------------------------------------
// predict_advection0 (runner.hpp:2609-2639)
template <bool W>
__global__ __launch_bounds__(256) void k_ii_predict0(SceneTab T, DevArrays A) {
  ALV_SCENE(T, A, s);
  if (s >= p.first + p.n) return;
  const float dt = A.ctl->dt;
  add_v(A, s, dt * ld3(A.a[s]));
  A.last_pressure[s] = A.pressure[s] * 0.5f;
  const vec3 xi = ld3(A.x[s]);
  const float dens = A.density[s];
  const float dens2 = dens * dens;
  vec3 d = mk(0.f, 0.f, 0.f);
  const unsigned nn = A.nnb[s];
  for (unsigned t0 = 0; t0 < nn; t0 += kNB) {
    unsigned j[kNB];
    batch_ids(A, s, nn, t0, j);
    float4 xj[kNB];
#pragma unroll
    for (int u = 0; u < kNB; ++u) xj[u] = A.x[j[u]];
#pragma unroll
    for (int u = 0; u < kNB; ++u) {
      if (t0 + u < nn) {
        vec3 xixj = wrapd<W>(xi - ld3(xj[u]), p);
        d -= (p.vol / dens2) * (gradW(xixj, p));
      }
    }
  }
  unsigned mask = A.bmask[s];
  while (mask) {
    const unsigned b = __builtin_ctz(mask);
    mask &= mask - 1u;
    float4 k = A.bker[size_t(b) * A.cap + s];
    d -= mk(k.x / dens2, k.y / dens2, k.z / dens2);
  }
  A.dii[s] = f4(d * p.rho0, 0.f);
}
// predict_advection1 (runner.hpp:2640-2690)
template <bool W>
__global__ __launch_bounds__(256) void k_ii_predict1(SceneTab T, DevArrays A) {
  ALV_SCENE(T, A, s);
  if (s >= p.first + p.n) return;
  const float dt = A.ctl->dt;
  const vec3 xi = ld3(A.x[s]), vi = ld3(A.v[s]), d_ii = ld3(A.dii[s]);
  const float dens = A.density[s];
  const float dens2 = dens * dens;
  const float inv_dens2 = 1.0f / dens2;
  const float dpi = p.vol * inv_dens2;
  float dadv = dens / p.rho0;
  float a_ii = 0.f;
  const unsigned nn = A.nnb[s];
  for (unsigned t0 = 0; t0 < nn; t0 += kNB) {
    unsigned j[kNB];
    batch_ids(A, s, nn, t0, j);
    float4 xj[kNB], vj[kNB];
#pragma unroll
    for (int u = 0; u < kNB; ++u) {
      xj[u] = A.x[j[u]];
      vj[u] = A.v[j[u]];
    }
#pragma unroll
    for (int u = 0; u < kNB; ++u) {
      if (t0 + u < nn) {
        vec3 xixj = wrapd<W>(xi - ld3(xj[u]), p);
        vec3 gw = gradW(xixj, p);
        vec3 dji = (dpi * p.rho0) * gw;
        dadv += dt * p.vol * dot(vi - ld3(vj[u]), gw);
        a_ii += p.vol * dot(d_ii - dji, gw);
      }
    }
  }
  unsigned mask = A.bmask[s];
  while (mask) {
    const unsigned b = __builtin_ctz(mask);
    mask &= mask - 1u;
    const size_t o = size_t(b) * A.cap + s;
    float4 r = A.brec[o], k = A.bker[o];
    vec3 g = mk(k.x, k.y, k.z);
    vec3 velj = rigid_velocity(A.bnd[b], ld3(r));
    vec3 dji = (inv_dens2 * p.rho0) * g;
    dadv += dt * dot(vi - velj, g);
    a_ii += dot(d_ii - dji, g);
  }
  A.adv_density[s] = dadv;
  A.aii[s] = a_ii * p.rho0;
}
// pressure_solve_iteration0 (runner.hpp:2691-2715).  The previous
// iteration's copy last_pressure <- pressure (solver_ii.hpp:204) is folded in:
// from the second iteration on, neighbour values are read from `pressure`
// and each particle refreshes its own last_pressure.
template <bool W>
__global__ __launch_bounds__(256) void k_ii_iter0(SceneTab T, DevArrays A) {
  ALV_SCENE(T, A, s);
  if (A.ctl->dens_done) return;
  if (s >= p.first + p.n) return;
  const float* lp = (A.ctl->dens_iter == 0u) ? A.last_pressure : A.pressure;
  const vec3 xi = ld3(A.x[s]);
  vec3 acc = mk(0.f, 0.f, 0.f);
  const unsigned nn = A.nnb[s];
  for (unsigned t0 = 0; t0 < nn; t0 += kNB) {
    unsigned j[kNB];
    batch_ids(A, s, nn, t0, j);
    float4 xj[kNB];
    float djs[kNB], lpj[kNB];
#pragma unroll
    for (int u = 0; u < kNB; ++u) {
      xj[u] = A.x[j[u]];
      djs[u] = A.density[j[u]];
      lpj[u] = lp[j[u]];
    }
#pragma unroll
    for (int u = 0; u < kNB; ++u) {
      if (t0 + u < nn) {
        vec3 xixj = wrapd<W>(xi - ld3(xj[u]), p);
        float dj = djs[u];
        float dj2 = dj * dj;
        acc -= (p.vol / dj2 * lpj[u]) * (gradW(xixj, p));
      }
    }
  }
  A.dij_pj[s] = f4(acc * p.rho0, 0.f);
  A.last_pressure[s] = lp[s];
}
// pressure_solve_iteration1 (runner.hpp:2716-2772) + FP64 partial sums
template <bool W>
__global__ __launch_bounds__(256) void k_ii_iter1(SceneTab T, DevArrays A) {
  __shared__ double lds[4];
  ALV_SCENE(T, A, s);
  if (A.ctl->dens_done) return;
  const float dt = A.ctl->dt;
  double part = 0.0;
  if (s < p.first + p.n) {
    const vec3 xi = ld3(A.x[s]);
    const float lp = A.last_pressure[s];
    const vec3 dpj = ld3(A.dij_pj[s]);
    const float dens = A.density[s];
    const float dens2 = dens * dens;
    const float dpi = p.vol / dens2;
    const float bb = 1.0f - A.adv_density[s];
    const float dt2 = dt * dt;
    const float a_ii = A.aii[s];
    const float denom = a_ii * dt2;
    const float w = 2.0f / 3.0f;
    float pr = 0.f;
    float sum_tmp = 0.f;
    const unsigned nn = A.nnb[s];
    for (unsigned t0 = 0; t0 < nn; t0 += kNB) {
      unsigned j[kNB];
      batch_ids(A, s, nn, t0, j);
      float4 xj[kNB], dpk[kNB], diij[kNB];
      float lpj[kNB];
#pragma unroll
      for (int u = 0; u < kNB; ++u) {
        xj[u] = A.x[j[u]];
        dpk[u] = A.dij_pj[j[u]];
        diij[u] = A.dii[j[u]];
        lpj[u] = A.last_pressure[j[u]];
      }
#pragma unroll
      for (int u = 0; u < kNB; ++u) {
        if (t0 + u < nn) {
          vec3 xixj = wrapd<W>(xi - ld3(xj[u]), p);
          vec3 djk_pk = ld3(dpk[u]);
          vec3 gw = gradW(xixj, p);
          vec3 dji = (dpi * p.rho0) * gw;
          vec3 dji_pi = dji * lp;
          sum_tmp += p.vol * dot(dpj - ld3(diij[u]) * lpj[u] - (djk_pk - dji_pi), gw);
        }
      }
    }
    unsigned mask = A.bmask[s];
    while (mask) {
      const unsigned b = __builtin_ctz(mask);
      mask &= mask - 1u;
      float4 k = A.bker[size_t(b) * A.cap + s];
      sum_tmp += dot(dpj, mk(k.x, k.y, k.z));
    }
    sum_tmp *= p.rho0;
    if (fabsf(denom) > 1.0e-9f)
      pr = std_max((1.0f - w) * lp + w / denom * (bb - dt2 * sum_tmp), 0.0f);
    float err = pr == 0.f ? 0.f : (denom * pr + sum_tmp * dt2 - bb);
    A.density_err[s] = err;
    A.pressure[s] = pr;
    part = double(err);
  }
  double t = block_sum_d(part, lds);
  if (threadIdx.x == 0) A.partials[xcd_block()] = t;
}
__global__ __launch_bounds__(kFinBlock) void k_ii_finalize(SceneTab T, DevArrays A) {
  __shared__ double lds[kFinBlock / 64];
  ALV_SCENE_BLOCK(T, A);
  StepCtl* ctl = A.ctl;
  if (ctl->dens_done) return;
  double part = 0.0;
  for (unsigned b = blk_begin + threadIdx.x; b < blk_end; b += blockDim.x) part += A.partials[b];
  double t = block_sum_d(part, lds);
  if (threadIdx.x == 0) {
    float mean = float(t / double(p.n));
    unsigned it = ctl->dens_iter + 1u;
    ctl->dens_iter = it;
    ctl->mean_density_error = mean;
    bool again = (mean > ctl->tol_dens || it < ctl->min_dens) && it < ctl->max_dens;
    ctl->dens_done = again ? 0u : 1u;
    publish_ctl(A, ctl, scene_);
  }
}
// compute_pressure_accels (runner.hpp:2774-2817)
template <bool W, bool PL>
__global__ __launch_bounds__(256) void k_ii_pressure_accel(SceneTab T, DevArrays A) {
  ALV_SCENE(T, A, s);
  if (s >= p.first + p.n) return;
  const vec3 xi = ld3(A.x[s]);
  const float dens = A.density[s];
  const float dens2 = dens * dens;
  const float dpi = A.pressure[s] / dens2;
  vec3 ai = mk(0.f, 0.f, 0.f);
  const unsigned nn = A.nnb[s];
  for (unsigned t0 = 0; t0 < nn; t0 += kNB) {
    unsigned j[kNB];
    batch_ids(A, s, nn, t0, j);
    float4 xj[kNB];
    float djs[kNB], pj[kNB];
#pragma unroll
    for (int u = 0; u < kNB; ++u) {
      xj[u] = A.x[j[u]];
      djs[u] = A.density[j[u]];
      pj[u] = A.pressure[j[u]];
    }
#pragma unroll
    for (int u = 0; u < kNB; ++u) {
      if (t0 + u < nn) {
        vec3 xixj = wrapd<W>(xi - ld3(xj[u]), p);
        float dj = djs[u];
        float dj2 = dj * dj;
        float dpj = pj[u] / dj2;
        ai -= (p.mass * (dpi + dpj)) * (gradW(xixj, p));
      }
    }
  }
  if constexpr (PL) {
    // compute_pressure_accels_with_pellets (runner.hpp:2819-2870)
    const unsigned npn = A.pnnb[s];
    for (unsigned t = 0; t < npn; ++t) {
      const unsigned sp = pellet_at(A.pnbr, A, s, t);
      const vec3 xixj = wrapd<W>(xi - ld3(A.px[sp]), p);
      const float dpj = A.pp[sp] / (p.rho0 * p.rho0);
      const vec3 acc = (p.mass * (dpi + dpj)) * gradW(xixj, p);
      ai -= acc;
      const unsigned b = A.pbody[sp];
      const BoundaryDev& B = A.bnd[b];
      if (B.accum) {
        const size_t o = size_t(b) * A.cap + s;
        const vec3 f = p.mass * acc;
        const vec3 tq = cross((xi - xixj) - mk(B.x[0], B.x[1], B.x[2]), f);
        float4 F = A.bforce[o], T = A.btorque[o];
        A.bforce[o] = f4(ld3(F) + f, 0.f);
        A.btorque[o] = f4(ld3(T) + tq, 0.f);
      }
    }
  }
  unsigned mask = PL ? 0u : A.bmask[s];
  while (mask) {
    const unsigned b = __builtin_ctz(mask);
    mask &= mask - 1u;
    const size_t o = size_t(b) * A.cap + s;
    float4 k = A.bker[o];
    vec3 acc = (p.rho0 * dpi) * mk(k.x, k.y, k.z);
    ai -= acc;
    const BoundaryDev& B = A.bnd[b];
    if (B.accum) {
      vec3 f = p.mass * acc;
      vec3 tq = cross(ld3(A.brec[o]) - mk(B.x[0], B.x[1], B.x[2]), f);
      float4 F = A.bforce[o], T = A.btorque[o];
      A.bforce[o] = f4(ld3(F) + f, 0.f);
      A.btorque[o] = f4(ld3(T) + tq, 0.f);
    }
  }
  A.pacc[s] = f4(p.rho0 * ai, 0.f);
}
// kinematic_integration<wrap> (runner.hpp:2873-2887) + back to user order
template <bool W>
__global__ __launch_bounds__(256) void k_ii_finish(SceneTab T, DevArrays A) {
  ALV_SCENE(T, A, s);
  if (s >= p.first + p.n) return;
  const float dt = A.ctl->dt;
  const unsigned i = A.sorted_id[s];
  vec3 v = ld3(A.v[s]);
  v += ld3(A.pacc[s]) * dt;
  vec3 x = wrapd<W>(ld3(A.x[s]) + v * dt, p);
  st3(A.ux, i, x);
  st3(A.uv, i, v);
  A.upressure[i] = A.pressure[s];
}

// ---------------------------------------------------------------------------
// ISPH: pressure-Laplacian Jacobi solve (SolverI::step, solver_i.hpp:79-330;
// kernels runner.hpp:2890-3057).  The Jacobi iterations ping-pong between
// last_pressure and pressure: iteration k reads buffer k & 1 and writes the
// other, which is the reference's "iterate, then last_pressure <- pressure"
// without the copy.  k_is_settle leaves the final values in `pressure`.
// ---------------------------------------------------------------------------
// advect_and_init_pressure (runner.hpp:2890-2900)
__global__ __launch_bounds__(256) void k_is_advect(SceneTab T, DevArrays A) {
  ALV_SCENE(T, A, s);
  if (s >= p.first + p.n) return;
  const float dt = A.ctl->dt;
  add_v(A, s, dt * ld3(A.a[s]));
  A.last_pressure[s] = A.pressure[s] * 0.5f;
}
// calculate_isph_diagonal_adv_density (runner.hpp:2901-2949): the Jacobi
// diagonal (times rho0) into aii, the advected density into adv_density
template <bool W, bool PL>
__global__ __launch_bounds__(256) void k_is_diag(SceneTab T, DevArrays A) {
  ALV_SCENE(T, A, s);
  if (s >= p.first + p.n) return;
  const float dt = A.ctl->dt;
  const vec3 xi = ld3(A.x[s]), vi = ld3(A.v[s]);
  const float dens = A.density[s];
  const float eps = 0.01f * p.h2;
  float dadv = dens / p.rho0;
  float diag = 0.f;
  const unsigned nn = A.nnb[s];
  for (unsigned t0 = 0; t0 < nn; t0 += kNB) {
    unsigned j[kNB];
    batch_ids(A, s, nn, t0, j);
    float4 xj[kNB], vj[kNB];
    float djs[kNB];
#pragma unroll
    for (int u = 0; u < kNB; ++u) {
      xj[u] = A.x[j[u]];
      vj[u] = A.v[j[u]];
      djs[u] = A.density[j[u]];
    }
#pragma unroll
    for (int u = 0; u < kNB; ++u) {
      if (t0 + u < nn) {
        const vec3 xixj = wrapd<W>(xi - ld3(xj[u]), p);
        const vec3 gw = gradW(xixj, p);
        const float dj = djs[u];
        dadv += dt * p.vol * dot(vi - ld3(vj[u]), gw);
        diag += p.vol * (dens + dj) / (dens * dj) / (length_sqr(xixj) + eps) * dot(xixj, gw);
      }
    }
  }
  if constexpr (PL) {
    // calculate_isph_diagonal_adv_density_with_pellets (runner.hpp:2951-3001)
    const unsigned npn = A.pnnb[s];
    for (unsigned t = 0; t < npn; ++t) {
      const unsigned sp = pellet_at(A.pnbr, A, s, t);
      const vec3 xixj = wrapd<W>(xi - ld3(A.px[sp]), p);
      const vec3 g = p.vol * gradW(xixj, p);
      dadv += dt * dot(vi - ld3(A.pv[sp]), g);
      diag += (dens + p.rho0) / (dens * p.rho0) / (length_sqr(xixj) + eps) * dot(xixj, g);
    }
  } else {
    unsigned mask = A.bmask[s];
    while (mask) {
      const unsigned b = __builtin_ctz(mask);
      mask &= mask - 1u;
      const size_t o = size_t(b) * A.cap + s;
      const float4 r = A.brec[o], k = A.bker[o];
      const vec3 g = mk(k.x, k.y, k.z);
      const vec3 xb = ld3(r);
      const vec3 xib = xi - xb;
      dadv += dt * dot(vi - rigid_velocity(A.bnd[b], xb), g);
      diag += (dens + p.rho0) / (dens * p.rho0) / (length_sqr(xib) + eps) * dot(xib, g);
    }
  }
  A.aii[s] = diag * p.rho0;
  A.adv_density[s] = dadv;
}
// isph_solve_iteration (runner.hpp:2996-3042) + FP64 partial sums
template <bool W, bool PL>
__global__ __launch_bounds__(256) void k_is_iter(SceneTab T, DevArrays A) {
  __shared__ double lds[4];
  ALV_SCENE(T, A, s);
  if (A.ctl->dens_done) return;
  const float dt = A.ctl->dt;
  const bool odd = (A.ctl->dens_iter & 1u) != 0u;
  const float* lp = odd ? A.pressure : A.last_pressure;
  float* out = odd ? A.last_pressure : A.pressure;
  double part = 0.0;
  if (s < p.first + p.n) {
    const vec3 xi = ld3(A.x[s]);
    const float dens = A.density[s];
    const float diag = A.aii[s];
    const float bb = 1.0f - A.adv_density[s];
    const float dt2 = dt * dt;
    const float denom = diag * dt2;
    const float eps = 0.01f * p.h2;
    const float w = 2.0f / 3.0f;
    float off = 0.f;
    const unsigned nn = A.nnb[s];
    for (unsigned t0 = 0; t0 < nn; t0 += kNB) {
      unsigned j[kNB];
      batch_ids(A, s, nn, t0, j);
      float4 xj[kNB];
      float djs[kNB], lpj[kNB];
#pragma unroll
      for (int u = 0; u < kNB; ++u) {
        xj[u] = A.x[j[u]];
        djs[u] = A.density[j[u]];
        lpj[u] = lp[j[u]];
      }
#pragma unroll
      for (int u = 0; u < kNB; ++u) {
        if (t0 + u < nn) {
          const vec3 xixj = wrapd<W>(xi - ld3(xj[u]), p);
          const float dj = djs[u];
          off -= p.vol * (dens + dj) / (dens * dj) / (length_sqr(xixj) + eps) *
                 dot(xixj, gradW(xixj, p)) * lpj[u];
        }
      }
    }
    if constexpr (PL) {
      // isph_solve_iteration_with_pellets (runner.hpp:3055-3118): pellets
      // at rest density with the extrapolated pressure
      const unsigned npn = A.pnnb[s];
      for (unsigned t = 0; t < npn; ++t) {
        const unsigned sp = pellet_at(A.pnbr, A, s, t);
        const vec3 xixj = wrapd<W>(xi - ld3(A.px[sp]), p);
        off -= p.vol * (dens + p.rho0) / (dens * p.rho0) / (length_sqr(xixj) + eps) *
               dot(xixj, gradW(xixj, p)) * A.pp[sp];
      }
    }
    off *= p.rho0;
    float pr = 0.f;
    if (fabsf(denom) > 1.0e-9f)
      pr = std_max((1.0f - w) * lp[s] + w * (bb / denom - off / diag), 0.0f);
    const float err = pr == 0.f ? 0.f : (denom * pr + off * dt2 - bb);
    A.density_err[s] = err;
    out[s] = pr;
    part = double(err);
  }
  double t = block_sum_d(part, lds);
  if (threadIdx.x == 0) A.partials[xcd_block()] = t;
}
// After an even, non-zero number of iterations the result is in
// last_pressure; with none, `pressure` keeps the previous step's values, as
// in the reference.
__global__ __launch_bounds__(256) void k_is_settle(SceneTab T, DevArrays A) {
  ALV_SCENE(T, A, s);
  if (s >= p.first + p.n) return;
  const unsigned it = A.ctl->dens_iter;
  if (it != 0u && (it & 1u) == 0u) A.pressure[s] = A.last_pressure[s];
}

// ---------------------------------------------------------------------------
// Rigid-body sums (solver_df.hpp:347-353): FP64, deterministic two-level
// reduction over the (particle, body) pairs with Psi != 0.  Grid (G, B):
// block (g, b) writes the partial sums of body b over its particle range.
// ---------------------------------------------------------------------------
constexpr unsigned kBodyBlocks = 256;

__global__ __launch_bounds__(256) void k_body_partials(SceneTab T, DevArrays A) {
  __shared__ double lds[4];
  const SimParams p = T.par[0];
  const unsigned b = blockIdx.y;
  if (!A.bnd[b].dynamic) return;
  const unsigned bit = 1u << b;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (unsigned s = blockIdx.x * kBlock + threadIdx.x; s < p.n; s += gridDim.x * kBlock) {
    if (p.pellet_mode || (A.bmask[s] & bit)) {
      const size_t o = size_t(b) * A.cap + s;
      float4 F = A.bforce[o], T = A.btorque[o];
      acc[0] += F.x;
      acc[1] += F.y;
      acc[2] += F.z;
      acc[3] += T.x;
      acc[4] += T.y;
      acc[5] += T.z;
    }
  }
  for (int c = 0; c < 6; ++c) {
    double t = block_sum_d(acc[c], lds);
    if (threadIdx.x == 0) A.body_partials[(size_t(b) * gridDim.x + blockIdx.x) * 6 + c] = t;
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_body_sums(SceneTab T, DevArrays A, unsigned nblk) {
  __shared__ double lds[4];
  const SimParams p = T.par[0];
  const unsigned b = blockIdx.x;
  if (!A.bnd[b].dynamic) return;
  for (int c = 0; c < 6; ++c) {
    double acc = 0.0;
    for (unsigned k = threadIdx.x; k < nblk; k += blockDim.x)
      acc += A.body_partials[(size_t(b) * nblk + k) * 6 + c];
    double t = block_sum_d(acc, lds);
    if (threadIdx.x == 0) A.body_sums[b * 6 + c] = t;  // host memory (mapped)
    __syncthreads();
  }
  if (threadIdx.x == 0) __threadfence_system();
}

// sorted -> user order for the API arrays that are synchronised lazily
__global__ __launch_bounds__(256) void k_unsort(SceneTab T, DevArrays A, unsigned mask) {
  ALV_SCENE(T, A, s);
  if (s >= p.first + p.n) return;
  const unsigned i = A.sorted_id[s];
  if (mask & ALV_F_DENSITY) A.udensity[i] = A.density[s];
  if (mask & ALV_F_ALPHA) A.ualpha[i] = A.alpha[s];
  if (mask & ALV_F_RHO_ADV) A.urho_adv[i] = A.rho_adv[s];
  if (mask & ALV_F_A) st3(A.ua, i, ld3(A.a[s]));
  if (mask & ALV_F_NNB) A.unnb[i] = A.nnb[s];
}

__global__ __launch_bounds__(256) void k_export_neighbors(SceneTab T, DevArrays A,
                                                          unsigned* ids, unsigned* counts) {
  ALV_SCENE(T, A, s);
  if (s >= p.first + p.n) return;
  const unsigned i = A.sorted_id[s];
  const unsigned nn = A.nnb[s];
  counts[i] = nn;
  for (unsigned t = 0; t < nn; ++t) {
    const uint4 g = A.nbr[nbr_group(A, t >> 2, s)];
    const unsigned j = (t & 3u) == 0 ? g.x : (t & 3u) == 1 ? g.y : (t & 3u) == 2 ? g.z : g.w;
    ids[size_t(i) * p.kmax + t] = A.sorted_id[j];
  }
}

// The neighbour list in the reference's layout (make_neighbor_list,
// runner.hpp:1896-1899): row i (user order) holds (x_i - x_j, bits of j)
// per neighbour, from the positions of the last neighbour search (the
// sorted positions keep them until the next search); slots past the count
// are zero.
template <bool W>
__global__ __launch_bounds__(256) void k_export_neighbor_rows(SceneTab T, DevArrays A,
                                                              float4* __restrict__ out) {
  ALV_SCENE(T, A, s);
  if (s >= p.first + p.n) return;
  const unsigned i = A.sorted_id[s];
  const unsigned nn = A.nnb[s];
  const vec3 xi = ld3(A.x[s]);
  float4* row = out + size_t(i) * p.kmax;
  for (unsigned t = 0; t < p.kmax; ++t) {
    if (t >= nn) {
      row[t] = make_float4(0.f, 0.f, 0.f, 0.f);
      continue;
    }
    const uint4 g = A.nbr[nbr_group(A, t >> 2, s)];
    const unsigned j = (t & 3u) == 0 ? g.x : (t & 3u) == 1 ? g.y : (t & 3u) == 2 ? g.z : g.w;
    const vec3 d = wrapd<W>(xi - ld3(A.x[j]), p);
    row[t] = make_float4(d.x, d.y, d.z, __uint_as_float(A.sorted_id[j]));
  }
}

// Boundary records and per-particle boundary forces in the reference's
// [B][max_num_particles] layout and user order (particle_boundary,
// particle_boundary_kernel: float4; particle_force, particle_torque:
// float3).  The step skips the pairs whose volume is 0; their records are
// computed here in full (closest boundary point, volume 0, kernel terms 0)
// at the positions and boundary kinematics of the last neighbour search, as
// the reference's boundary pass writes them for every pair
// (compute_particle_boundary_analytic, runner.hpp:1982-2002).  Forces of
// boundaries that accumulate none are zero.  Null outputs are skipped.
__global__ __launch_bounds__(256) void k_export_boundary(SceneTab T, DevArrays A,
                                                         float4* __restrict__ rec,
                                                         float4* __restrict__ ker,
                                                         float* __restrict__ force,
                                                         float* __restrict__ torque) {
  ALV_SCENE(T, A, s);
  if (s >= p.first + p.n) return;
  const unsigned i = A.sorted_id[s];
  const unsigned mask = p.pellet_mode ? 0u : A.bmask[s];
  const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
  for (unsigned b = 0; b < p.num_boundaries; ++b) {
    const size_t o = size_t(b) * A.cap + s;
    const size_t u = size_t(b) * A.cap + i;
    const bool seen = (mask >> b) & 1u;
    if (rec || ker) {
      float4 r = zero, k = zero;
      if (seen) {
        r = A.brec[o];
        k = A.bker[o];
      } else if (!p.pellet_mode) {
        boundary_record<false>(A.bnd[b], ld3(A.x[s]), p, r, k);
      }
      if (rec) rec[u] = r;
      if (ker) ker[u] = k;
    }
    const bool has_force = A.bnd[b].accum && (seen || p.pellet_mode);
    if (force) {
      const float4 F = has_force ? A.bforce[o] : zero;
      force[3 * u] = F.x;
      force[3 * u + 1] = F.y;
      force[3 * u + 2] = F.z;
    }
    if (torque) {
      const float4 Q = has_force ? A.btorque[o] : zero;
      torque[3 * u] = Q.x;
      torque[3 * u + 1] = Q.y;
      torque[3 * u + 2] = Q.z;
    }
  }
}

// IISPH per-particle arrays in user order (SolverIi's particle_last_pressure,
// particle_aii, particle_adv_density, particle_density_err; particle_dii,
// particle_dij_pj, particle_pressure_accel as float3).  Null outputs are
// skipped.
__global__ __launch_bounds__(256) void k_export_iisph(SceneTab T, DevArrays A, float* last_p,
                                                      float* aii, float* adv, float* err,
                                                      float* dii, float* dij, float* pacc) {
  ALV_SCENE(T, A, s);
  if (s >= p.first + p.n) return;
  const unsigned i = A.sorted_id[s];
  if (last_p) last_p[i] = A.last_pressure[s];
  if (aii) aii[i] = A.aii[s];
  if (adv) adv[i] = A.adv_density[s];
  if (err) err[i] = A.density_err[s];
  if (dii) st3(dii, i, ld3(A.dii[s]));
  if (dij) st3(dij, i, ld3(A.dij_pj[s]));
  if (pacc) st3(pacc, i, ld3(A.pacc[s]));
}

// ---------------------------------------------------------------------------
// Volume maps (update_volume_field, runner.hpp:1616-1694).  Peirce constants
// are packed as cosphi[30], sinphi[30], b[30], r[15], c[15].
// ---------------------------------------------------------------------------
constexpr float kPiF = 3.14159265358979323846f;

__global__ __launch_bounds__(256) void k_volume_analytic(SimParams p, BoundaryDev B,
                                                         float* volume, const float* pc,
                                                         unsigned begin, unsigned end) {
  const unsigned l = begin + blockIdx.x * kBlock + threadIdx.x;
  if (l >= end) return;
  const float* cosphi = pc;
  const float* sinphi = pc + 30;
  const float* kb = pc + 60;
  const float* kr = pc + 90;
  const float* kc = pc + 105;
  vec3 x = node_position(B.geom, l);
  float sum = 0.0f;
  for (unsigned i = 0; i < 60; ++i) {
    float theta = 2.0f * kPiF * float(i + 1) / 60.0f;
    float ct = cosf(theta), st = sinf(theta);
    for (unsigned j = 0; j < 30; ++j) {
      for (unsigned k = 0; k < 15; ++k) {
        float wijk = 2.0f * kPiF * kb[j] * kc[k] / 60.0f;
        vec3 u = mk(sinphi[j] * ct, sinphi[j] * st, cosphi[j]);
        vec3 q = (p.h * kr[k]) * u;
        float dist = sdf_eval(B, x + q);
        float dm = dist * B.sign;
        sum += wijk * (dm <= 0.0f ? 1.0f : W_dist(dm, p) / p.w0);
      }
    }
  }
  volume[l] = sum * p.h2 * p.h;
}

__global__ __launch_bounds__(256) void k_volume_grid(SimParams p, GridGeomDev g,
                                                     const float* __restrict__ dist,
                                                     float* volume, float sign,
                                                     const float* pc, unsigned begin,
                                                     unsigned end) {
  const unsigned l = begin + blockIdx.x * kBlock + threadIdx.x;
  if (l >= end) return;
  const float* cosphi = pc;
  const float* sinphi = pc + 30;
  const float* kb = pc + 60;
  const float* kr = pc + 90;
  const float* kc = pc + 105;
  vec3 x = node_position(g, l);
  float sum = 0.0f;
  const float wi = 2.0f * kPiF / 60.0f;
  for (unsigned i = 0; i < 60; ++i) {
    float theta = 2.0f * kPiF * float(i + 1) / 60.0f;
    float ct = cosf(theta), st = sinf(theta);
    for (unsigned j = 0; j < 30; ++j) {
      vec3 u = mk(sinphi[j] * ct, sinphi[j] * st, cosphi[j]);
      float wij = wi * kb[j];
      for (unsigned k = 0; k < 15; ++k) {
        float wijk = wij * kc[k];
        vec3 q = (p.h * kr[k]) * u;
        float d = grid_sample(g, dist, x + q, FLT_MAX);
        float dm = d * sign;
        sum += wijk * (dm <= 0.0f ? 1.0f : W_dist(dm, p) / p.w0);
      }
    }
  }
  volume[l] = sum * p.h2 * p.h;
}

__global__ __launch_bounds__(256) void k_node_positions(GridGeomDev g, float* out) {
  const unsigned l = blockIdx.x * kBlock + threadIdx.x;
  if (l >= g.num_nodes) return;
  vec3 x = node_position(g, l);
  out[3 * l] = x.x;
  out[3 * l + 1] = x.y;
  out[3 * l + 2] = x.z;
}

// ---------------------------------------------------------------------------
// Particle-sampled boundaries: pellets (Pile with VolumeMethod::pellets,
// SolverI::step with pellets, solver_i.hpp:79-330).  The pellets of every
// body are placed at the step start, sorted into a cell grid of their own
// (the fluid's cells), and the fluid particles keep a second list of the
// pellets around them; in the reference both lists come from one grid of
// fluid and pellets, where pellet ids follow every fluid id, so each list
// in (cell, id) order is the same.
// ---------------------------------------------------------------------------
// The 27 cells around x in the reference order (dx outer, dz inner; periodic
// axes wrapped), each in ascending sorted index: f(j, x - x_j) for every j
// of `xs` closer than the kernel radius.
template <bool W, class F>
__device__ inline void scan_cells(const SimParams& p, const unsigned* cell_start,
                                  const float4* xs, vec3 x, F&& f) {
  int ip[3];
  cell_ipos(x, p, ip);
  for (int t = 0; t < 27; ++t) {
    int q[3] = {ip[0] + t / 9 - 1, ip[1] + (t / 3) % 3 - 1, ip[2] + t % 3 - 1};
    if constexpr (W) {
      for (int a = 0; a < 3; ++a)
        if (p.wrap_mask & (1 << a))
          q[a] = q[a] >= p.res[a] ? q[a] - p.res[a] : (q[a] < 0 ? q[a] + p.res[a] : q[a]);
    }
    if (!within_grid(q, p)) continue;
    const unsigned c = p.cell_base + cell_linear(q, p);
    const unsigned e = cell_start[c + 1];
    for (unsigned j = cell_start[c]; j < e; ++j) {
      const vec3 d = wrapd<W>(x - ld3(xs[j]), p);
      if (length_sqr(d) < p.h2) f(j, d);
    }
  }
}
__device__ inline void list_put(uint4* list, const DevArrays& A, unsigned s, unsigned t,
                                unsigned v) {
  reinterpret_cast<unsigned*>(list + nbr_group(A, t >> 2, s))[t & 3u] = v;
}

// transform_pellets (runner.hpp:4071-4085) with the step-start kinematics:
// pellet k of the pile into ux/uv row cap + k.  One thread per pellet.
__global__ __launch_bounds__(256) void k_transform_pellets(SceneTab T, DevArrays A) {
  ALV_SCENE_LINEAR(T, A, k);
  if (k >= p.n) return;
  const BoundaryDev& B = A.bnd[A.pvbody[k]];
  const vec3 r = rotate(ld3(A.pverts, k), make_float4(B.q[0], B.q[1], B.q[2], B.q[3]));
  st3(A.ux, A.cap + k, r + mk(B.x[0], B.x[1], B.x[2]));
  st3(A.uv, A.cap + k,
      cross(mk(B.omega[0], B.omega[1], B.omega[2]), r) + mk(B.v[0], B.v[1], B.v[2]));
}
// user -> sorted order of the pellets
__global__ __launch_bounds__(256) void k_pellet_gather(SceneTab T, DevArrays A) {
  ALV_SCENE_LINEAR(T, A, s);
  if (s >= p.n) return;
  const unsigned k = A.psorted_id[s];
  A.px[s] = f4(ld3(A.ux, A.cap + k), 0.f);
  A.pv[s] = f4(ld3(A.uv, A.cap + k), 0.f);
  A.pbody[s] = A.pvbody[k];
}
// The pellet lists of the fluid particles (make_bead_pellet_neighbor_list,
// runner.hpp:1913-1971; the fluid lists are k_neighbors') and the density
// with pellets (compute_density_with_pellets, runner.hpp:2074-2098):
// rho = m (W(0) + sum_fluid W + sum_pellets W), both sums over the lists.
template <bool W>
__global__ __launch_bounds__(256) void k_pellet_fluid_lists(SceneTab T, DevArrays A) {
  ALV_SCENE(T, A, s);
  if (s >= p.first + p.n) return;
  const unsigned K = p.kmax;
  const vec3 xi = ld3(A.x[s]);
  unsigned cnt = 0;
  scan_cells<W>(p, A.pcell_start, A.px, xi, [&](unsigned j, vec3) {
    list_put(A.pnbr, A, s, min(cnt, K - 1), j);
    ++cnt;
  });
  const unsigned npn = min(cnt, K);
  A.pnnb[s] = npn;
  if (cnt > K) A.anomaly[2] = 1u;
  float dens = p.w0;
  const unsigned nn = A.nnb[s];
  for (unsigned t = 0; t < nn; ++t)
    dens += W_d2(length_sqr(wrapd<W>(xi - ld3(A.x[pellet_at(A.nbr, A, s, t)]), p)), p);
  for (unsigned t = 0; t < npn; ++t)
    dens += W_d2(length_sqr(wrapd<W>(xi - ld3(A.px[pellet_at(A.pnbr, A, s, t)]), p)), p);
  const float rho = dens * p.mass;
  A.density[s] = rho;
  reinterpret_cast<float*>(A.v + s)[3] = visc_coefficient(p, rho);
}
// The fluid lists of the pellets (the bead lists of the same search), which
// the MLS extrapolation reads.  One thread per pellet.
template <bool W>
__global__ __launch_bounds__(256) void k_pellet_bead_lists(SceneTab T, DevArrays A) {
  ALV_SCENE_LINEAR(T, A, s);
  if (s >= p.n) return;
  const unsigned K = p.kmax;
  unsigned cnt = 0;
  scan_cells<W>(p, A.cell_start, A.x, ld3(A.px[s]), [&](unsigned j, vec3) {
    list_put(A.pbnbr, A, s, min(cnt, K - 1), j);
    ++cnt;
  });
  A.pbnnb[s] = min(cnt, K);
  if (cnt > K) A.anomaly[2] = 1u;
}
// compute_viscosity_with_pellets (runner.hpp:2251-2297) with
// clear_acceleration and the per-block CFL maxima, as k_viscosity
template <bool W>
__global__ __launch_bounds__(256) void k_viscosity_pellets(SceneTab T, DevArrays A) {
  __shared__ float lds[4];
  ALV_SCENE(T, A, s);
  const float dt = A.ctl->dt;
  float cfl = -FLT_MAX;
  if (s < p.first + p.n) {
    const vec3 xi = ld3(A.x[s]), vi = ld3(A.v[s]);
    vec3 da = viscosity_sum<W>(p, A, s, xi, vi, A.nnb[s], GlobalXV{A.x, A.v});
    const unsigned npn = A.pnnb[s];
    for (unsigned t = 0; t < npn; ++t) {
      const unsigned sp = pellet_at(A.pnbr, A, s, t);
      const vec3 xixj = wrapd<W>(xi - ld3(A.px[sp]), p);
      const vec3 acc = (p.vol * p.bvisc * dot(xixj, gradW(xixj, p)) /
                        (length_sqr(xixj) + 0.01f * p.h2)) *
                       (vi - ld3(A.pv[sp]));
      da += acc;
      const unsigned b = A.pbody[sp];
      const BoundaryDev& B = A.bnd[b];
      if (B.accum) {
        const size_t o = size_t(b) * A.cap + s;
        const vec3 f = (-p.mass) * acc;
        const vec3 tq = cross((xi - xixj) - mk(B.x[0], B.x[1], B.x[2]), f);
        float4 F = A.bforce[o], Tq = A.btorque[o];
        A.bforce[o] = f4(ld3(F) + f, 0.f);
        A.btorque[o] = f4(ld3(Tq) + tq, 0.f);
      }
    }
    const vec3 a = mk(p.g[0], p.g[1], p.g[2]) + da;
    A.a[s] = f4(a, 0.f);
    cfl = length_sqr(vi + a * dt);
  }
  float m = block_max_f(cfl, lds);
  if (threadIdx.x == 0) A.fpartials[xcd_block()] = m;
}

// Pseudo-inverse of the symmetric 3x3 moment matrix of the MLS plane fit
// (runner.hpp:2581-2590 takes it from an SVD, runner.hpp:340-574).  For a
// symmetric matrix V S^+ U^T equals sum_k v_k v_k^T / lambda_k over the
// eigenpairs with |lambda_k| > 1e-2, the reference's singular-value cut.
// Eigenpairs by six sweeps of cyclic Jacobi rotations (Jacobi 1846; the
// rotation of Golub and Van Loan, Matrix Computations, 8.5.2).
__device__ inline void sym_pinv(vec3 d, vec3 o, float inv[3][3]) {
  float m[3][3] = {{d.x, o.x, o.y}, {o.x, d.y, o.z}, {o.y, o.z, d.z}};
  float e[3][3] = {{1.f, 0.f, 0.f}, {0.f, 1.f, 0.f}, {0.f, 0.f, 1.f}};
#pragma unroll
  for (int sweep = 0; sweep < 6; ++sweep) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const int pp = r < 2 ? 0 : 1, qq = r == 0 ? 1 : 2, k = 3 - pp - qq;
      const float apq = m[pp][qq];
      if (apq == 0.f) continue;
      const float theta = (m[qq][qq] - m[pp][pp]) / (2.f * apq);
      const float tn = (theta >= 0.f ? 1.f : -1.f) / (fabsf(theta) + sqrtf(theta * theta + 1.f));
      const float cs = 1.f / sqrtf(tn * tn + 1.f), sn = tn * cs;
      m[pp][pp] -= tn * apq;
      m[qq][qq] += tn * apq;
      m[pp][qq] = m[qq][pp] = 0.f;
      const float akp = m[k][pp], akq = m[k][qq];
      m[k][pp] = m[pp][k] = cs * akp - sn * akq;
      m[k][qq] = m[qq][k] = sn * akp + cs * akq;
#pragma unroll
      for (int w = 0; w < 3; ++w) {
        const float ep = e[w][pp], eq = e[w][qq];
        e[w][pp] = cs * ep - sn * eq;
        e[w][qq] = sn * ep + cs * eq;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) inv[r][c] = 0.f;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float lam = m[k][k];
    if (!(fabsf(lam) > 1e-2f)) continue;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) inv[r][c] += e[r][k] * e[c][k] / lam;
  }
}
// extrapolate_pressure_to_pellet (runner.hpp:2543-2607): the weighted plane
// fit of the fluid neighbours' pressure at each pellet, clamped at 0 (fmax:
// a pellet without fluid neighbours gets 0).  Inside the Jacobi loop it
// reads the buffer the next iteration reads; final_pressure reads the
// solved pressure and also stores the pellets' pressure in user order.
template <bool W>
__global__ __launch_bounds__(256) void k_extrapolate(SceneTab T, DevArrays A, int final_pressure) {
  ALV_SCENE_LINEAR(T, A, s);
  if (s >= p.n) return;
  if (!final_pressure && A.ctl->dens_done) return;
  const float* src = (final_pressure || (A.ctl->dens_iter & 1u) != 0u) ? A.pressure
                                                                        : A.last_pressure;
  const vec3 xi = ld3(A.px[s]);
  const unsigned nn = A.pbnnb[s];
  float w_sum = 0.f;
  vec3 center = mk(0.f, 0.f, 0.f);
  for (unsigned t = 0; t < nn; ++t) {
    const vec3 xj = ld3(A.x[pellet_at(A.pbnbr, A, s, t)]);
    const float w = W_d2(length_sqr(wrapd<W>(xi - xj, p)), p) * p.vol;
    w_sum += w;
    center += w * xj;
  }
  center = mk(center.x / w_sum, center.y / w_sum, center.z / w_sum);
  float alpha = 0.f;
  vec3 dg = mk(0.f, 0.f, 0.f), od = mk(0.f, 0.f, 0.f), pxs = mk(0.f, 0.f, 0.f);
  for (unsigned t = 0; t < nn; ++t) {
    const unsigned j = pellet_at(A.pbnbr, A, s, t);
    const vec3 xj = ld3(A.x[j]);
    const vec3 xt = xj - center;
    const float w = W_d2(length_sqr(wrapd<W>(xi - xj, p)), p) * p.vol;
    const float wp = w * src[j];
    alpha += wp;
    pxs += wp * xt;
    dg += w * mk(xt.x * xt.x, xt.y * xt.y, xt.z * xt.z);
    od += w * mk(xt.x * xt.y, xt.x * xt.z, xt.y * xt.z);
  }
  alpha /= w_sum;
  float inv[3][3];
  sym_pinv(dg, od, inv);
  const vec3 plane = mk(inv[0][0] * pxs.x + inv[0][1] * pxs.y + inv[0][2] * pxs.z,
                        inv[1][0] * pxs.x + inv[1][1] * pxs.y + inv[1][2] * pxs.z,
                        inv[2][0] * pxs.x + inv[2][1] * pxs.y + inv[2][2] * pxs.z);
  const float pr = fmaxf(alpha + dot(plane, xi - center), 0.f);
  A.pp[s] = pr;
  if (final_pressure) A.upressure[A.cap + A.psorted_id[s]] = pr;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
KernelProfiler* g_profiler = nullptr;

#define ALV_LAUNCH_GRID(kernel, grid, s, ...)               \
  do {                                                      \
    if (g_profiler) g_profiler->begin(#kernel, s);          \
    hipLaunchKernelGGL(kernel, grid, dim3(kBlock), 0, s, __VA_ARGS__); \
    if (g_profiler) g_profiler->end(#kernel, s);            \
  } while (0)
// one thread per particle slot of every scene
#define ALV_LAUNCH(kernel, L, s, ...) ALV_LAUNCH_GRID(kernel, dim3((L).pblocks), s, (L).T, __VA_ARGS__)
// kernels templated on the periodic wrap and one more flag
#define ALV_LAUNCH_W2(kernel, flag, L, s, ...)                                    \
  do {                                                                            \
    if (g_profiler) g_profiler->begin(#kernel, s);                                \
    if ((L).p0.wrap_mask)                                                         \
      hipLaunchKernelGGL((kernel<true, flag>), dim3((L).pblocks), dim3(kBlock), 0, s, (L).T, \
                         __VA_ARGS__);                                            \
    else                                                                          \
      hipLaunchKernelGGL((kernel<false, flag>), dim3((L).pblocks), dim3(kBlock), 0, s, (L).T, \
                         __VA_ARGS__);                                            \
    if (g_profiler) g_profiler->end(#kernel, s);                                  \
  } while (0)
// neighbour passes with an LDS-tile instance (never with the periodic wrap):
// the gather instances reserve no LDS, so they keep their occupancy
#define ALV_LAUNCH_WT(kernel, L, s, ...)                                          \
  do {                                                                            \
    if (g_profiler) g_profiler->begin(#kernel, s);                                \
    if ((L).p0.wrap_mask)                                                         \
      hipLaunchKernelGGL((kernel<true, false>), dim3((L).pblocks), dim3(kBlock), 0, s, (L).T, \
                         __VA_ARGS__);                                            \
    else if ((L).tiles)                                                           \
      hipLaunchKernelGGL((kernel<false, true>), dim3((L).pblocks), dim3(kBlock), 0, s, (L).T, \
                         __VA_ARGS__);                                            \
    else                                                                          \
      hipLaunchKernelGGL((kernel<false, false>), dim3((L).pblocks), dim3(kBlock), 0, s, (L).T, \
                         __VA_ARGS__);                                            \
    if (g_profiler) g_profiler->end(#kernel, s);                                  \
  } while (0)
// density-change passes (ALV_SCENE_PASS): one block per particle block, of
// 256 threads for the tile instance and kLPP * 256 for the others
#define ALV_PASS_BLOCK(tiled) dim3((tiled) ? kBlock : kLPP * kBlock)
#define ALV_LAUNCH_WTP(kernel, L, s, ...)                                         \
  do {                                                                            \
    if (g_profiler) g_profiler->begin(#kernel, s);                                \
    if ((L).p0.wrap_mask)                                                         \
      hipLaunchKernelGGL((kernel<true, false>), dim3((L).pblocks), ALV_PASS_BLOCK(false), 0, s, \
                         (L).T, __VA_ARGS__);                                     \
    else if ((L).tiles)                                                           \
      hipLaunchKernelGGL((kernel<false, true>), dim3((L).pblocks), ALV_PASS_BLOCK(true), 0, s, \
                         (L).T, __VA_ARGS__);                                     \
    else                                                                          \
      hipLaunchKernelGGL((kernel<false, false>), dim3((L).pblocks), ALV_PASS_BLOCK(false), 0, s, \
                         (L).T, __VA_ARGS__);                                     \
    if (g_profiler) g_profiler->end(#kernel, s);                                  \
  } while (0)
#define ALV_LAUNCH_W2TP(kernel, flag, L, s, ...)                                  \
  do {                                                                            \
    if (g_profiler) g_profiler->begin(#kernel, s);                                \
    if ((L).p0.wrap_mask)                                                         \
      hipLaunchKernelGGL((kernel<true, flag, false>), dim3((L).pblocks), ALV_PASS_BLOCK(false), 0, \
                         s, (L).T, __VA_ARGS__);                                  \
    else if ((L).tiles)                                                           \
      hipLaunchKernelGGL((kernel<false, flag, true>), dim3((L).pblocks), ALV_PASS_BLOCK(true), 0, \
                         s, (L).T, __VA_ARGS__);                                  \
    else                                                                          \
      hipLaunchKernelGGL((kernel<false, flag, false>), dim3((L).pblocks), ALV_PASS_BLOCK(false), \
                         0, s, (L).T, __VA_ARGS__);                               \
    if (g_profiler) g_profiler->end(#kernel, s);                                  \
  } while (0)
#define ALV_LAUNCH_W2T(kernel, flag, L, s, ...)                                   \
  do {                                                                            \
    if (g_profiler) g_profiler->begin(#kernel, s);                                \
    if ((L).p0.wrap_mask)                                                         \
      hipLaunchKernelGGL((kernel<true, flag, false>), dim3((L).pblocks), dim3(kBlock), 0, s, \
                         (L).T, __VA_ARGS__);                                     \
    else if ((L).tiles)                                                           \
      hipLaunchKernelGGL((kernel<false, flag, true>), dim3((L).pblocks), dim3(kBlock), 0, s, \
                         (L).T, __VA_ARGS__);                                     \
    else                                                                          \
      hipLaunchKernelGGL((kernel<false, flag, false>), dim3((L).pblocks), dim3(kBlock), 0, s, \
                         (L).T, __VA_ARGS__);                                     \
    if (g_profiler) g_profiler->end(#kernel, s);                                  \
  } while (0)
// kernels templated on the periodic wrap: one name for both instances
#define ALV_LAUNCH_W(kernel, L, s, ...)                                           \
  do {                                                                            \
    if (g_profiler) g_profiler->begin(#kernel, s);                                \
    if ((L).p0.wrap_mask)                                                         \
      hipLaunchKernelGGL(kernel<true>, dim3((L).pblocks), dim3(kBlock), 0, s, (L).T, __VA_ARGS__); \
    else                                                                          \
      hipLaunchKernelGGL(kernel<false>, dim3((L).pblocks), dim3(kBlock), 0, s, (L).T, __VA_ARGS__); \
    if (g_profiler) g_profiler->end(#kernel, s);                                  \
  } while (0)
// one block per scene
#define ALV_LAUNCH1(kernel, s, ...) ALV_LAUNCH_GRID(kernel, dim3(1), s, __VA_ARGS__)
// one block of kFinBlock threads per scene
#define ALV_LAUNCH_FIN(kernel, L, s, ...)                                          \
  do {                                                                             \
    if (g_profiler) g_profiler->begin(#kernel, s);                                 \
    hipLaunchKernelGGL(kernel, dim3((L).scenes), dim3(kFinBlock), 0, s, (L).T, __VA_ARGS__); \
    if (g_profiler) g_profiler->end(#kernel, s);                                   \
  } while (0)

__global__ __launch_bounds__(256) void k_fill_u32(unsigned* __restrict__ d, unsigned v,
                                                  size_t n) {
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) d[i] = v;
}
// zeroes the cell counts and the three grid anomaly flags and resets every
// scene's control block (one launch)
__global__ __launch_bounds__(256) void k_clear_grid(DevArrays A, unsigned cells, unsigned scenes) {
  const unsigned i = blockIdx.x * kBlock + threadIdx.x;
  if (i < cells) A.cell_count[i] = 0u;
  if (i < 3u) A.anomaly[i] = 0u;
  if (i < scenes) reset_ctl(A, i);
}
void launch_clear_grid(const Launch& L, const DevArrays& A, hipStream_t s) {
  const unsigned nb = std::max(1u, (std::max(L.cells, L.scenes) + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_clear_grid, dim3(nb), dim3(kBlock), 0, s, A, L.cells, L.scenes);
}
void launch_fill_u32(unsigned* dst, unsigned value, size_t count, hipStream_t s) {
  if (!count) return;
  hipLaunchKernelGGL(k_fill_u32, dim3(unsigned((count + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     s, dst, value, count);
}

// The slots of a scene's range past its particles are counted into its
// outside-the-grid bin, so that after the scan every scene's sorted range
// starts where its user range starts.
__global__ void k_reserve_padding(SceneTab T, DevArrays A, unsigned scenes) {
  const unsigned sc = blockIdx.x * blockDim.x + threadIdx.x;
  if (sc >= scenes) return;
  const SimParams p = T.par[sc];
  A.cell_count[p.cell_base + p.num_cells] += p.span - p.n;
}
void launch_reserve_padding(const Launch& L, const DevArrays& A, hipStream_t s) {
  hipLaunchKernelGGL(k_reserve_padding, dim3((L.scenes + 63) / 64), dim3(64), 0, s, L.T, A,
                     L.scenes);
}

void launch_cell_count(const Launch& L, const DevArrays& A, hipStream_t s) {
  ALV_LAUNCH(k_cell_count, L, s, A);
}
void launch_scan(unsigned count, const unsigned* in, unsigned* out, unsigned* tmp,
                 hipStream_t s) {
  const unsigned nb = (count + kScanPerBlock - 1) / kScanPerBlock;
  ALV_LAUNCH_GRID(k_scan_reduce, dim3(nb), s, count, in, tmp);
  ALV_LAUNCH1(k_scan_top, s, nb, tmp);
  ALV_LAUNCH_GRID(k_scan_apply, dim3(nb), s, count, in, out, (const unsigned*)tmp);
}
void launch_scatter(const Launch& L, const DevArrays& A, hipStream_t s) {
  ALV_LAUNCH(k_scatter, L, s, A);
}
void launch_cell_sort(const Launch& L, const DevArrays& A, hipStream_t s) {
  ALV_LAUNCH(k_cell_sort, L, s, A);
}
void launch_tiles(const Launch& L, const DevArrays& A, hipStream_t s) {
  if (!A.use_tiles) return;
  ALV_LAUNCH_GRID(k_tiles, dim3(grid_for(L.pblocks)), s, L.T, A, L.pblocks);
}
void launch_gather_boundary(const Launch& L, const DevArrays& A, int kind, hipStream_t s) {
  ALV_LAUNCH(k_gather_boundary, L, s, A, kind);
}
void launch_neighbors(const Launch& L, const DevArrays& A, int with_factor, hipStream_t s) {
  ALV_LAUNCH_WT(k_neighbors, L, s, A, with_factor);
}
void launch_df_warm_start_div(const Launch& L, const DevArrays& A, hipStream_t s,
                              int features, const MlpDev* net, int net_divergence, int fused) {
  if (!fused) {  // else warm start 0 ran in the neighbour search (mode 2)
    if (net && features)
      ALV_LAUNCH_W2(k_ws_div0_learned, true, L, s, A, *net, net_divergence);
    else if (net)
      ALV_LAUNCH_W2(k_ws_div0_learned, false, L, s, A, *net, net_divergence);
    else if (features)
      ALV_LAUNCH_W2TP(k_ws_div0, true, L, s, A);
    else
      ALV_LAUNCH_W2TP(k_ws_div0, false, L, s, A);
  }
  ALV_LAUNCH_WT(k_ws_div1, L, s, A);
  ALV_LAUNCH_WTP(k_vel_density_change, L, s, A);
}
void launch_df_div_iteration(const Launch& L, const DevArrays& A, hipStream_t s) {
  ALV_LAUNCH_WT(k_div_iter, L, s, A);
  ALV_LAUNCH_WTP(k_div_err, L, s, A);
  ALV_LAUNCH_FIN(k_div_finalize, L, s, A);
}
void launch_df_viscosity(const Launch& L, const DevArrays& A, hipStream_t s) {
  ALV_LAUNCH_WT(k_viscosity, L, s, A);
}
void launch_update_dt(const Launch& L, const DevArrays& A, hipStream_t s) {
  ALV_LAUNCH_FIN(k_update_dt, L, s, A);
}
void launch_df_integrate_np(const Launch& L, const DevArrays& A, int fin, hipStream_t s) {
  ALV_LAUNCH(k_integrate_np, L, s, A, fin);
}
void launch_df_warm_start_dens(const Launch& L, const DevArrays& A, hipStream_t s,
                               int learned) {
  if (learned)
    ALV_LAUNCH_W2TP(k_ws_p0, true, L, s, A);
  else
    ALV_LAUNCH_W2TP(k_ws_p0, false, L, s, A);
  ALV_LAUNCH_WT(k_ws_p1, L, s, A);
  ALV_LAUNCH_WTP(k_rho_adv, L, s, A);
}
void launch_df_dens_iteration(const Launch& L, const DevArrays& A, hipStream_t s) {
  ALV_LAUNCH_WT(k_p_iter, L, s, A);
  ALV_LAUNCH_WTP(k_p_err, L, s, A);
  ALV_LAUNCH_FIN(k_p_finalize, L, s, A);
}
void launch_df_finish(const Launch& L, const DevArrays& A, int dens_finish, hipStream_t s) {
  ALV_LAUNCH_W(k_df_finish, L, s, A, dens_finish);
}
void launch_export_pairs(const Launch& L, const DevArrays& A, float* x, float* y,
                         hipStream_t s) {
  ALV_LAUNCH(k_export_pairs, L, s, A, x, y);
}
void launch_density_change(const Launch& L, const DevArrays& A, hipStream_t s) {
  ALV_LAUNCH_WTP(k_density_change, L, s, A);
}
void launch_integrate_velocity(const Launch& L, const DevArrays& A, float dt, hipStream_t s) {
  ALV_LAUNCH_W(k_integrate_velocity, L, s, A, dt);
}
void launch_ii_viscosity_predict(const Launch& L, const DevArrays& A, hipStream_t s) {
  ALV_LAUNCH_WT(k_viscosity, L, s, A);
}
void launch_ii_predict(const Launch& L, const DevArrays& A, hipStream_t s) {
  ALV_LAUNCH_W(k_ii_predict0, L, s, A);
  ALV_LAUNCH_W(k_ii_predict1, L, s, A);
}
void launch_ii_iteration(const Launch& L, const DevArrays& A, hipStream_t s) {
  ALV_LAUNCH_W(k_ii_iter0, L, s, A);
  ALV_LAUNCH_W(k_ii_iter1, L, s, A);
  ALV_LAUNCH_FIN(k_ii_finalize, L, s, A);
}
void launch_is_prepare(const Launch& L, const DevArrays& A, hipStream_t s) {
  ALV_LAUNCH(k_is_advect, L, s, A);
  ALV_LAUNCH_W2(k_is_diag, false, L, s, A);
}
void launch_is_iteration(const Launch& L, const DevArrays& A, hipStream_t s) {
  ALV_LAUNCH_W2(k_is_iter, false, L, s, A);
  ALV_LAUNCH_FIN(k_ii_finalize, L, s, A);
}
void launch_is_settle(const Launch& L, const DevArrays& A, hipStream_t s) {
  ALV_LAUNCH(k_is_settle, L, s, A);
}
void launch_ii_finish(const Launch& L, const DevArrays& A, hipStream_t s) {
  ALV_LAUNCH_W2(k_ii_pressure_accel, false, L, s, A);
  ALV_LAUNCH_W(k_ii_finish, L, s, A);
}
void launch_body_sums(const Launch& L, const DevArrays& A, hipStream_t s) {
  const SimParams& p = L.p0;
  if (p.num_boundaries == 0) return;
  const unsigned nblk = std::min(kBodyBlocks, std::max(1u, grid_for(p.n)));
  ALV_LAUNCH_GRID(k_body_partials, dim3(nblk, p.num_boundaries), s, L.T, A);
  ALV_LAUNCH_GRID(k_body_sums, dim3(p.num_boundaries), s, L.T, A, nblk);
}
void launch_unsort_fields(const Launch& L, const DevArrays& A, unsigned mask, hipStream_t s) {
  ALV_LAUNCH(k_unsort, L, s, A, mask);
}
void launch_export_neighbors(const Launch& L, const DevArrays& A, unsigned* ids,
                             unsigned* counts, hipStream_t s) {
  ALV_LAUNCH(k_export_neighbors, L, s, A, ids, counts);
}
void launch_export_neighbor_rows(const Launch& L, const DevArrays& A, float4* out,
                                 hipStream_t s) {
  ALV_LAUNCH_W(k_export_neighbor_rows, L, s, A, out);
}
void launch_export_boundary(const Launch& L, const DevArrays& A, float4* rec, float4* ker,
                            float* force, float* torque, hipStream_t s) {
  ALV_LAUNCH(k_export_boundary, L, s, A, rec, ker, force, torque);
}
void launch_export_iisph(const Launch& L, const DevArrays& A, float* const scalar[4],
                         float* const vec[3], hipStream_t s) {
  ALV_LAUNCH(k_export_iisph, L, s, A, scalar[0], scalar[1], scalar[2], scalar[3], vec[0],
             vec[1], vec[2]);
}
void launch_volume_field_analytic(const SimParams& p, const BoundaryDev& b, float* volume,
                                  const float* peirce, unsigned begin, unsigned end,
                                  hipStream_t s) {
  if (end <= begin) return;
  hipLaunchKernelGGL(k_volume_analytic, dim3(grid_for(end - begin)), dim3(kBlock), 0, s, p, b,
                     volume, peirce, begin, end);
}
void launch_volume_field_grid(const SimParams& p, const GridGeomDev& g, const float* distance,
                              float* volume, float sign, const float* peirce, unsigned begin,
                              unsigned end, hipStream_t s) {
  if (end <= begin) return;
  hipLaunchKernelGGL(k_volume_grid, dim3(grid_for(end - begin)), dim3(kBlock), 0, s, p, g,
                     distance, volume, sign, peirce, begin, end);
}
// The device arithmetic the kernel functions are built on, element-wise
// (alv_check_math): 0 sqrt_rn(a), 1 div_known(a, b, RN(1/b)), 2 gradW_scale
// of the displacement (a, 0, 0), 3 W_d2(a).
__global__ __launch_bounds__(256) void k_check_math(SimParams p, int op, const float* a,
                                                    const float* b, float* out, size_t n) {
  const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const float x = a[i];
  float r = 0.0f;
  switch (op) {
    case 0: r = sqrt_rn(x); break;
    case 1: r = div_known(x, b[i], 1.0f / b[i]); break;
    case 2: r = gradW_scale<true>(mk(x, 0.0f, 0.0f), p); break;
    case 6: r = gradW_scale<false>(mk(x, 0.0f, 0.0f), p); break;
    case 3: r = W_d2(x, p); break;
    case 4: r = rcp_rn(x); break;
    case 5: r = div_rn(x, b[i]); break;
    default: break;
  }
  out[i] = r;
}
void launch_check_math(const SimParams& p, int op, const float* a, const float* b, float* out,
                       size_t n, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_check_math, dim3(unsigned((n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     s, p, op, a, b, out, n);
}
void launch_node_positions(const GridGeomDev& g, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_node_positions, dim3(grid_for(g.num_nodes)), dim3(kBlock), 0, s, g, out);
}

void launch_transform_pellets(const Launch& Lp, const DevArrays& A, hipStream_t s) {
  if (Lp.active) ALV_LAUNCH_GRID(k_transform_pellets, dim3(Lp.pblocks), s, Lp.T, A);
}
void launch_pellet_grid(const Launch& Lp, const DevArrays& A, hipStream_t s) {
  // the pellets' view of the counting sort: their user rows, keys and grid
  DevArrays Ap = A;
  Ap.ux = A.ux + size_t(A.cap) * 3;
  Ap.uv = A.uv + size_t(A.cap) * 3;
  Ap.key = A.pkey;
  Ap.slot = A.pslot;
  Ap.sorted_id = A.psorted_id;
  Ap.sort_tmp = A.psort_tmp;
  Ap.cell_count = A.pcell_count;
  Ap.cell_start = A.pcell_start;
  Ap.scan_tmp = A.pscan_tmp;
  launch_fill_u32(A.pcell_count, 0u, Lp.cells, s);
  if (Lp.active) launch_cell_count(Lp, Ap, s);
  launch_reserve_padding(Lp, Ap, s);
  launch_scan(Lp.cells, Ap.cell_count, Ap.cell_start, Ap.scan_tmp, s);
  if (Lp.active) {
    launch_scatter(Lp, Ap, s);
    launch_cell_sort(Lp, Ap, s);
    ALV_LAUNCH_GRID(k_pellet_gather, dim3(Lp.pblocks), s, Lp.T, A);
  }
}
void launch_pellet_lists(const Launch& L, const Launch& Lp, const DevArrays& A, hipStream_t s) {
  if (L.active) ALV_LAUNCH_W(k_pellet_fluid_lists, L, s, A);
  if (Lp.active) ALV_LAUNCH_W(k_pellet_bead_lists, Lp, s, A);
}
void launch_pellet_viscosity(const Launch& L, const DevArrays& A, hipStream_t s) {
  ALV_LAUNCH_W(k_viscosity_pellets, L, s, A);
}
void launch_is_prepare_pellets(const Launch& L, const DevArrays& A, hipStream_t s) {
  ALV_LAUNCH(k_is_advect, L, s, A);
  ALV_LAUNCH_W2(k_is_diag, true, L, s, A);
}
void launch_is_iteration_pellets(const Launch& L, const Launch& Lp, const DevArrays& A,
                                 hipStream_t s) {
  if (Lp.active) ALV_LAUNCH_W(k_extrapolate, Lp, s, A, 0);
  ALV_LAUNCH_W2(k_is_iter, true, L, s, A);
  ALV_LAUNCH_FIN(k_ii_finalize, L, s, A);
}
void launch_is_finish_pellets(const Launch& L, const Launch& Lp, const DevArrays& A,
                              hipStream_t s) {
  if (Lp.active) ALV_LAUNCH_W(k_extrapolate, Lp, s, A, 1);
  ALV_LAUNCH_W2(k_ii_pressure_accel, true, L, s, A);
  ALV_LAUNCH_W(k_ii_finish, L, s, A);
}

}  // namespace alv
