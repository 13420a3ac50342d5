// alv_kernels.hpp -- host-side launchers for the HIP kernels of the solver
// core.  All launches are asynchronous on the given stream.
#pragma once

#include <hip/hip_runtime.h>

#include "alv_mlp.hpp"
#include "alv_types.hpp"

namespace alv {

// LDS tile of one 256-particle block: the neighbours of a block whose
// particles all lie in one x-slab of its scene's grid are inside three
// contiguous ranges of sorted particles (the slab and the two beside it),
// which a pass copies into LDS with coalesced loads before the neighbour
// loops.  Ranges are in increasing order; absent ranges have lo = ~0u.
constexpr unsigned kTileMax = 2496;  // particles per tile (78 KB of x and v: two blocks per CU)
struct Tile {
  unsigned lo[3];
  unsigned len[3];
  unsigned total;
  unsigned ok;  // the block may use its tile
};

// Device pointer bundle of one solver (passed to kernels by value).
// "u*" arrays are in user (API) order and hold AoS float3 vectors like the
// reference's Variable<1,float3>; all other per-particle arrays are in
// cell-sorted order, SoA, float4-padded vectors.
struct DevArrays {
  unsigned cap;  // max_num_particles: row stride of [B, cap]
  unsigned kmax;  // neighbour slots per particle
  // user order
  float* ux;
  float* uv;
  float* ua;
  float* udensity;
  float* ualpha;
  float* ukappa;
  float* ukappa_v;
  float* urho_adv;
  unsigned* unnb;
  float* upressure;
  // counting sort
  unsigned* key;
  unsigned* slot;
  unsigned* sorted_id;  // sorted position -> user index
  unsigned* sort_tmp;   // the scatter's arrival-ordered user indices
  unsigned* cell_count;  // [num_cells + 1]; last bin = outside the grid
  unsigned* cell_start;  // [num_cells + 2]
  unsigned* scan_tmp;
  unsigned* anomaly;  // [3]
  // sorted order
  float4* x;
  float4* v;
  float4* a;
  float* density;
  float* alpha;
  float* kappa;
  float* kappa_v;
  float* rho_adv;
  float* kfac;  // stiffness k_j of the running Jacobi iteration
  Tile* tiles;    // [cap / 256] LDS tiles of the neighbour passes
  int use_tiles;  // passes stage neighbour tiles in LDS
  float4* feat;  // [2 cap] warm-start features (learned warm start, logging)
  float* kpred;  // predicted stored density stiffness (learned warm start)
  uint4* nbr;     // neighbour list: groups of 4 slots in wave tiles (nbr_group)
  unsigned* nnb;
  float4* brec;  // [B, cap] (x_b, Psi)
  float4* bker;  // [B, cap] (Psi gradW, Psi W)
  unsigned* bmask;  // bit b set when Psi_b != 0
  float4* bforce;  // [B, cap]
  float4* btorque;  // [B, cap]
  // IISPH (sorted order)
  float* pressure;
  float* last_pressure;
  float* aii;
  float4* dii;
  float4* dij_pj;
  float* adv_density;
  float* density_err;
  float4* pacc;
  // Particle-sampled boundaries (pellets).  The pile's pellets (body
  // frame, user order) are placed in world space into ux/uv rows
  // [cap, cap + np) at the step start and sorted by cell on a grid of their
  // own (same cells as the fluid's); lists hold sorted indices.
  unsigned pcap;  // max_num_pellets
  const float* pverts;  // [pcap][3] body-frame pellets (user order)
  const unsigned* pvbody;  // body of every pellet (user order)
  unsigned* pkey;
  unsigned* pslot;
  unsigned* psorted_id;  // sorted pellet -> user pellet index
  unsigned* psort_tmp;
  unsigned* pcell_count;
  unsigned* pcell_start;
  unsigned* pscan_tmp;
  float4* px;  // sorted pellet positions
  float4* pv;  // sorted pellet velocities
  unsigned* pbody;  // sorted pellet -> body
  float* pp;  // pellet pressure (MLS extrapolation), sorted
  uint4* pnbr;  // pellet lists of the fluid particles (nbr_group layout)
  unsigned* pnnb;
  uint4* pbnbr;  // fluid lists of the pellets (nbr_group layout over pellets)
  unsigned* pbnnb;
  // reductions and control
  double* partials;  // per-block FP64 partial sums
  float* fpartials;  // per-block float maxima
  double* body_partials;  // [B, 256, 6]
  // [B, 6] in the host's pinned, coherent buffer (mapped): k_body_sums
  // stores the sums there, so the host reads them without a copy command
  double* body_sums;
  StepCtl* ctl;
  // the host's pinned, coherent copy of the control blocks (mapped): every
  // kernel that changes a block publishes it there (publish_ctl)
  StepCtl* ctl_pub;
  const BoundaryDev* bnd;
};

constexpr unsigned kBlock = 256;
// Lanes per particle of the density-change passes that gather neighbours
// from global memory (2 or 4; the lanes of a particle lie in one quad).
constexpr unsigned kLPP = 2;
static_assert(kLPP == 2u || kLPP == 4u, "a particle's lanes lie in one quad");

// Neighbour slots are stored in groups of four (one uint4 per particle and
// group).  Each 64-particle wave owns one contiguous [groups][64] tile of
// uint4, so a wave reading group q of its particles reads 1 KB contiguous,
// and the search writes whole 16-byte groups (buffered in registers) rather
// than single slots.  Slot t of particle s is component t % 4 of group t / 4.
__host__ __device__ inline unsigned nbr_groups(unsigned kmax) { return (kmax + 3u) / 4u; }
__host__ __device__ inline size_t nbr_group(const DevArrays& A, unsigned q, unsigned s) {
  return (size_t(s >> 6) * nbr_groups(A.kmax) + q) * 64u + (s & 63u);
}
inline size_t nbr_capacity(unsigned kmax, unsigned cap) {
  return size_t(nbr_groups(kmax)) * ((size_t(cap) + 63u) / 64u * 64u);
}

// Optional per-kernel timing hook (HIP events around each launch); null
// when profiling is off.
struct KernelProfiler {
  virtual ~KernelProfiler() = default;
  virtual void begin(const char* name, hipStream_t s) = 0;
  virtual void end(const char* name, hipStream_t s) = 0;
};
extern KernelProfiler* g_profiler;
inline unsigned grid_for(unsigned n) { return (n + kBlock - 1) / kBlock; }

// Scene table of a solver (device memory): the parameters of every scene and
// the scene of every 256-particle block and every 256-cell block.  A solver
// of one scene has a table of one entry.
struct SceneTab {
  const SimParams* par;
  const unsigned short* blk_scene;
  const unsigned short* cblk_scene;
};

// Host-side description of one launch set.
struct Launch {
  SceneTab T;
  unsigned pblocks;  // particle blocks: sum of span / 256
  unsigned cblocks;  // cell blocks
  unsigned cells;    // cell slots: sum of the per-scene cell ranges
  unsigned scenes;
  unsigned active;   // particles in all scenes
  SimParams p0;      // scene 0 (shared constants, wrap flags)
  int tiles;         // launch the LDS-tile instances of the neighbour passes
};

// Particle-kernel prologue: this block's particle index s (blocks remapped
// per XCD, or linear), its scene's parameters p, and the scene's control
// block and boundaries in A.
#define ALV_SCENE_AT(T, A, s, blk)                       \
  const unsigned s = (blk) * kBlock + threadIdx.x;       \
  const unsigned scene_ = (T).blk_scene[(blk)];          \
  const SimParams p = (T).par[scene_];                   \
  (A).ctl += scene_;                                     \
  (A).bnd += p.bnd_base
#define ALV_SCENE(T, A, s) ALV_SCENE_AT(T, A, s, xcd_block())
// Prologue of the density-change passes: with ONE (their LDS-tile
// instances) one lane per particle as ALV_SCENE, else kLPP lanes per
// particle in blocks of kLPP * 256 threads (one block per particle block);
// q is the lane within the particle's lanes.
#define ALV_SCENE_PASS(T, A, s, q, ONE)                                               \
  constexpr unsigned lpp_ = (ONE) ? 1u : kLPP;                                        \
  const unsigned qblk_ = xcd_block();                                                 \
  const unsigned q = threadIdx.x & (lpp_ - 1u);                                       \
  const unsigned s = qblk_ * kBlock + threadIdx.x / lpp_;                             \
  const unsigned scene_ = (T).blk_scene[qblk_];                                       \
  const SimParams p = (T).par[scene_];                                                \
  (A).ctl += scene_;                                                                  \
  (A).bnd += p.bnd_base
#define ALV_SCENE_LINEAR(T, A, s) ALV_SCENE_AT(T, A, s, blockIdx.x)
// One block per scene (reductions over a scene's blocks).
#define ALV_SCENE_BLOCK(T, A)                            \
  const unsigned scene_ = blockIdx.x;                    \
  const SimParams p = (T).par[scene_];                   \
  (A).ctl += scene_;                                     \
  (A).bnd += p.bnd_base;                                 \
  const unsigned blk_begin = p.first / kBlock, blk_end = blk_begin + p.span / kBlock

// neighbour search (counting sort)
void launch_fill_u32(unsigned* dst, unsigned value, size_t count, hipStream_t s);
// zeroes the cell counts (L.cells) and the grid anomaly flags and resets the
// control blocks for the step's solves
void launch_clear_grid(const Launch& L, const DevArrays& A, hipStream_t s);
void launch_cell_count(const Launch& L, const DevArrays& A, hipStream_t s);
// reserves each scene's unused slots at the end of its outside-the-grid bin
void launch_reserve_padding(const Launch& L, const DevArrays& A, hipStream_t s);
// Cells per block of the cell-count scan (1024: four per thread, so a 1 M-cell
// grid runs 1024 blocks, four per CU); `tmp` holds one sum per block.
constexpr unsigned kScanPerBlock = 1024;
void launch_scan(unsigned count, const unsigned* in, unsigned* out, unsigned* tmp,
                 hipStream_t s);
void launch_scatter(const Launch& L, const DevArrays& A, hipStream_t s);
void launch_cell_sort(const Launch& L, const DevArrays& A, hipStream_t s);
// LDS tile ranges of every block (after the sort, before the passes)
void launch_tiles(const Launch& L, const DevArrays& A, hipStream_t s);
// gather user -> sorted and compute boundary records
void launch_gather_boundary(const Launch& L, const DevArrays& A, int solver_kind,
                            hipStream_t s);
// neighbour list + density; mode 1: + DFSPH factor; mode 2: + factor and the
// divergence solve's warm start 0
void launch_neighbors(const Launch& L, const DevArrays& A, int with_factor, hipStream_t s);

// DFSPH
// features: record the eight warm-start features; net: learned warm start
// (null = the reference's warm start)
// fused: warm start 0 already ran in the neighbour search (launch_neighbors
// mode 2)
void launch_df_warm_start_div(const Launch& L, const DevArrays& A, hipStream_t s,
                              int features, const MlpDev* net, int net_divergence, int fused);
void launch_df_div_iteration(const Launch& L, const DevArrays& A, hipStream_t s);
// the viscosity of a step without a divergence solve (with one, that solve's
// density-change passes compute it)
void launch_df_viscosity(const Launch& L, const DevArrays& A, hipStream_t s);
void launch_update_dt(const Launch& L, const DevArrays& A, hipStream_t s);
// fin: the divergence solve ran (and computed the viscosity): also finish it
// and add the bodies' viscous reactions
void launch_df_integrate_np(const Launch& L, const DevArrays& A, int fin, hipStream_t s);
void launch_df_warm_start_dens(const Launch& L, const DevArrays& A, hipStream_t s,
                               int learned);
// training pairs of the last step in sorted order: x [n][8], y [n][2]
void launch_export_pairs(const Launch& L, const DevArrays& A, float* x, float* y,
                         hipStream_t s);
void launch_df_dens_iteration(const Launch& L, const DevArrays& A, hipStream_t s);
void launch_df_finish(const Launch& L, const DevArrays& A, int dens_finish, hipStream_t s);
void launch_density_change(const Launch& L, const DevArrays& A, hipStream_t s);
void launch_integrate_velocity(const Launch& L, const DevArrays& A, float dt, hipStream_t s);

// IISPH
void launch_ii_viscosity_predict(const Launch& L, const DevArrays& A, hipStream_t s);
void launch_ii_predict(const Launch& L, const DevArrays& A, hipStream_t s);
void launch_ii_iteration(const Launch& L, const DevArrays& A, hipStream_t s);
// non-zero flags of the volume-map cells (alv_sample.hip)
void launch_map_flags(const GridGeomDev& g, const float* nodes, unsigned char* out,
                      hipStream_t s);
// sampling and statistics (alv_sample.hip)
void launch_sample_neighbors(const SimParams& p, const DevArrays& A, const float* sample_x,
                             unsigned ns, float4* out, unsigned* counts, hipStream_t s);
// make_bead_pellet_neighbor_list on samples offset .. offset+ns (rows
// indexed by the sample's particle index) against a pellet solver's grids
void launch_sample_bead_pellet(const SimParams& p, const DevArrays& A, const float* sample_x,
                               unsigned ns, unsigned offset, unsigned max_beads, float4* bead,
                               unsigned* bead_counts, float4* pel, unsigned* pel_counts,
                               hipStream_t s);
// sampled density (0), velocity (1) or vorticity (2) with pellet lists
void launch_sample_with_pellets(const SimParams& p, int what, const float4* nbr,
                                const unsigned* counts, const float4* pel,
                                const unsigned* pcounts, unsigned K, unsigned ns,
                                const float* density, const float* v, const float* sample_v,
                                float* out, hipStream_t s);
// what: 0 boundary kernel with pellets (float4 out), 1 density with
// pellets, 2 DFSPH factor with pellets
void launch_pellet_sums(const SimParams& p, int what, const float4* bead,
                        const unsigned* bead_counts, const float4* pel,
                        const unsigned* pel_counts, unsigned K, unsigned n, float* out,
                        hipStream_t s);
void launch_sample_boundary(const SimParams& p, const BoundaryDev* bodies, unsigned nb,
                            const float* sample_x, unsigned ns, float4* rec, float4* ker,
                            hipStream_t s);
void launch_sample_fluid(const SimParams& p, const float4* nbr, const unsigned* counts,
                         unsigned K, unsigned ns, const float* density, const float* q,
                         unsigned comps, float* out, hipStream_t s);
void launch_sample_density(const SimParams& p, const float4* nbr, const unsigned* counts,
                           unsigned K, unsigned ns, const float4* ker, unsigned nb, float* out,
                           hipStream_t s);
void launch_sample_velocity(const SimParams& p, const float4* nbr, const unsigned* counts,
                            unsigned K, unsigned ns, const float* density, const float* v,
                            const float* sample_v, const float4* rec, const float4* ker,
                            unsigned nb, const float* rx, const float* rv, const float* rw,
                            float* out, hipStream_t s);
void launch_histogram256(const float* v, const float* mask, unsigned n, float v_min,
                         float step_inverse, unsigned* partial, unsigned* hist, hipStream_t s);
unsigned histogram_partial_count();
// compute_boundary_mask (analytic): mask[i] *= (sign d(local x_i) >= threshold)
void launch_boundary_mask(const BoundaryDev& B, const float* x, float threshold, float* mask,
                          unsigned n, hipStream_t s);
void launch_kl_divergence(const unsigned* hp, const unsigned* hq, unsigned n_p, unsigned n_q,
                          float q_lower_bound, unsigned bins, float* out, hipStream_t s);

// SolverI (ISPH)
void launch_is_prepare(const Launch& L, const DevArrays& A, hipStream_t s);
void launch_is_iteration(const Launch& L, const DevArrays& A, hipStream_t s);
void launch_is_settle(const Launch& L, const DevArrays& A, hipStream_t s);
void launch_ii_finish(const Launch& L, const DevArrays& A, hipStream_t s);

// Pellets (Lp: the launch set of the pellets, one thread per pellet).
// transform_all_pellets: places the pellets (ux/uv rows cap + k) with the
// body kinematics of A.bnd
void launch_transform_pellets(const Launch& Lp, const DevArrays& A, hipStream_t s);
// sorts the pellets at their ux rows into their grid (empty when none)
void launch_pellet_grid(const Launch& Lp, const DevArrays& A, hipStream_t s);
// the fluid particles' pellet lists and density with pellets, then the
// pellets' fluid lists
void launch_pellet_lists(const Launch& L, const Launch& Lp, const DevArrays& A, hipStream_t s);
void launch_pellet_viscosity(const Launch& L, const DevArrays& A, hipStream_t s);
void launch_is_prepare_pellets(const Launch& L, const DevArrays& A, hipStream_t s);
void launch_is_iteration_pellets(const Launch& L, const Launch& Lp, const DevArrays& A,
                                 hipStream_t s);
// MLS pressure of the pellets from the final fluid pressure, the pressure
// accelerations with pellets and the integration
void launch_is_finish_pellets(const Launch& L, const Launch& Lp, const DevArrays& A,
                              hipStream_t s);

// rigid bodies: per-body FP64 force / torque sums into A.body_sums
void launch_body_sums(const Launch& L, const DevArrays& A, hipStream_t s);
// sorted -> user order for the lazily synchronised API arrays
void launch_unsort_fields(const Launch& L, const DevArrays& A, unsigned field_mask,
                          hipStream_t s);
// reference-layout neighbour ids (user order) for inspection: ids[i*kmax+t]
void launch_export_neighbors(const Launch& L, const DevArrays& A, unsigned* ids,
                             unsigned* counts, hipStream_t s);
// reference-layout neighbour rows (x_i - x_j, bits of j), [cap][kmax] float4
void launch_export_neighbor_rows(const Launch& L, const DevArrays& A, float4* out,
                                 hipStream_t s);
// boundary records, boundary kernels ([B][cap] float4) and per-particle
// boundary forces and torques ([B][cap] float3) in user order; null outputs
// are skipped
void launch_export_boundary(const Launch& L, const DevArrays& A, float4* rec, float4* ker,
                            float* force, float* torque, hipStream_t s);
// IISPH arrays in user order: scalar = last_pressure, aii, adv_density,
// density_err; vec = dii, dij_pj, pressure_accel (float3); nulls skipped
void launch_export_iisph(const Launch& L, const DevArrays& A, float* const scalar[4],
                         float* const vec[3], hipStream_t s);

// volume maps (setup)
void launch_volume_field_analytic(const SimParams& p, const BoundaryDev& b, float* volume,
                                  const float* peirce, unsigned begin, unsigned end,
                                  hipStream_t s);
void launch_volume_field_grid(const SimParams& p, const GridGeomDev& g,
                              const float* distance, float* volume, float sign,
                              const float* peirce, unsigned begin, unsigned end,
                              hipStream_t s);
void launch_node_positions(const GridGeomDev& g, float* out, hipStream_t s);
// element-wise device arithmetic of the kernel functions (alv_check_math)
void launch_check_math(const SimParams& p, int op, const float* a, const float* b, float* out,
                       size_t n, hipStream_t s);

// field sync masks for launch_unsort_fields
enum : unsigned {
  ALV_F_DENSITY = 1u << 0,
  ALV_F_ALPHA = 1u << 1,
  ALV_F_RHO_ADV = 1u << 2,
  ALV_F_A = 1u << 3,
  ALV_F_NNB = 1u << 4,
  ALV_F_PRESSURE = 1u << 5,
};

// field codes written by launch_unsort_fields / gathered in gather_boundary
// solver state gathered into sorted order by k_gather_boundary: DFSPH
// stiffnesses (with DF_FACTOR also the DFSPH factor, for a step without the
// divergence solve, which does not compute it), IISPH/ISPH pressure, or none
// (positions and velocities only)
enum { ALV_KIND_DF = 0, ALV_KIND_II = 1, ALV_KIND_NONE = 2, ALV_KIND_DF_FACTOR = 3 };

}  // namespace alv
