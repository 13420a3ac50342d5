/*
 * alv.h -- C ABI of the MI355X (gfx950) SPH solver core.
 *
 * This is the drop-in boundary for the per-step solver path of
 * kennychufk/alluvion.  The reference exposes that path as C++ templates
 * (Store, Variable, Runner<TF>, Pile<TF>, SolverDf<TF>, SolverIi<TF>) bound to
 * Python by pybind11 (lib/pyalluvion.cu:1279-1486).  Here the same operations
 * are plain extern "C" functions over opaque handles, raw device pointers and
 * sizes; the header-only C++ mirror alv/alluvion.hpp and the pybind11 module
 * alluvion_amd._alluvion are thin layers on top.  Every function returns
 * ALV_OK or an error code; alv_last_error() gives the message (the reference
 * throws std::runtime_error, allocator.cu:35-41).
 *
 * Precision: FP32 state and arithmetic; convergence and rigid-force sums are
 * accumulated in FP64.  Not thread-safe per handle; one HIP stream per solver.
 */
#ifndef ALV_ALV_H
#define ALV_ALV_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ALV_OK 0
#define ALV_ERR_HIP 1
#define ALV_ERR_ARG 2
#define ALV_ERR_UNSUPPORTED 3
#define ALV_ERR_STATE 4

/* Message of the last failing call on this thread. */
const char* alv_last_error(void);
/* Library version string. */
const char* alv_version(void);

/* ---- device and memory: Store / Variable / Allocator -------------------- */
/* Store::set_device (lib/store.cu:18-20) */
int alv_set_device(int device);
int alv_get_device_count(int* count);
int alv_device_synchronize(void);
/* Store::create / Store::remove device buffers (store.hpp:45-84,
 * allocator.hpp:13-24).  Memory is zero-filled. */
int alv_malloc(void** dptr, size_t bytes);
int alv_free(void* dptr);
/* Store::create_pinned (store.hpp:56-66) */
int alv_malloc_pinned(void** hptr, size_t bytes);
int alv_free_pinned(void* hptr);
/* Variable::set_bytes / get_bytes / set_from / set_zero
 * (variable.hpp:60-131) */
int alv_memcpy_h2d(void* dst, const void* src, size_t bytes);
int alv_memcpy_d2h(void* dst, const void* src, size_t bytes);
int alv_memcpy_d2d(void* dst, const void* src, size_t bytes);
int alv_memset(void* dptr, int value, size_t bytes);
/* Runner::sum / Runner::max over n FP32 primitives (runner.hpp:4646-4675);
 * the sum accumulates in FP64. */
int alv_sum_f32(const float* d, size_t n, double* out);
int alv_max_f32(const float* d, size_t n, float* out);
/* alv_sum_f32 without any allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (tight stepping loops
 * that must not allocate, harnesses that forbid hipMalloc).  The block
 * partials and the result live in a fixed device array of the library
 * (ALV_SUM_NOALLOC_SCRATCH_DOUBLES doubles, 8,200 bytes per GPU, held until
 * the process ends); the caller passes the input d (device, n floats) and
 * the host result *out.  Same kernels and summation order as alv_sum_f32, so
 * the same bits.  n above ALV_SUM_NOALLOC_MAX_ELEMENTS fails with ALV_ERR_ARG.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (several host threads) race.  Runs on the null stream and waits for
 * the result like alv_sum_f32. */
#define ALV_SUM_NOALLOC_SCRATCH_DOUBLES 1025
#define ALV_SUM_NOALLOC_MAX_ELEMENTS (1ull << 36)
int alv_sum_f32_noalloc(const float* d, size_t n, double* out);

/* ---- constants: Const<float> and Consti (constants.hpp:40-137) ----------- */
typedef struct alv_const {
  float kernel_radius;
  float kernel_radius_sqr;
  float cubic_kernel_k;
  float cubic_kernel_l;
  float cubic_kernel_zero;
  float particle_radius;
  float particle_vol;
  float particle_mass;
  float density0;
  float viscosity;
  float boundary_viscosity;
  float gravity[3];
  float boundary_epsilon;
  float dfsph_factor_epsilon;
  float contact_tolerance;
  /* periodic wrap; [1] is the reference's y-wrap (set_wrap_length), [0] and
   * [2] extend it to x and z; length 0 disables an axis */
  float wrap_length[3];
} alv_const;

typedef struct alv_consti {
  unsigned max_num_particles_per_cell;
  unsigned grid_res[3];
  int grid_offset[3];
  unsigned max_num_neighbors_per_particle;
  unsigned num_boundaries;
  unsigned max_num_contacts;
} alv_consti;

/* Const::set_kernel_radius / set_particle_attr / set_wrap_length
 * (constants.hpp:100-136), computed in FP32 like the reference. */
void alv_const_set_kernel_radius(alv_const* cn, float h);
void alv_const_set_particle_attr(alv_const* cn, float radius, float mass,
                                 float density);

/* ---- boundaries: Pile<float> (pile.hpp:20-1066) --------------------------- */
typedef struct alv_pile alv_pile;

/* Analytic distances (dg/box_distance.hpp, dg/box_shell_distance.hpp,
 * dg/sphere_distance.hpp), and a distance sampled on the boundary's map
 * nodes (ALV_DISTANCE_MAP: the reference's grid path, which it takes for
 * MeshDistance, pile.hpp:79-95,508-521, runner.hpp:1696-1736,2005-2025). */
#define ALV_DISTANCE_BOX 0
#define ALV_DISTANCE_BOX_SHELL 1
#define ALV_DISTANCE_SPHERE 2
#define ALV_DISTANCE_MAP 3
typedef struct alv_distance {
  int type;
  /* box: widths[3], outset; box shell: inner widths[3], thickness, outset;
   * sphere: radius */
  float params[5];
  /* ALV_DISTANCE_MAP: the distance's bounding box (body frame), and a
   * function that writes the signed distance of n points (xyz triples, body
   * frame) to out[n].  Pile::add calls it on the nodes of the boundary's map
   * (in chunks, before it returns), as construct_distance_grid evaluates the
   * mesh distance; the map is then interpolated with its gradient.  Unused
   * for the analytic types. */
  float aabb_min[3], aabb_max[3];
  void (*signed_distance)(void* context, const float* points, unsigned n, float* out);
  void* context;
} alv_distance;

/* VolumeMethod (pile.hpp:16): volume maps, or particle-sampled boundaries
 * ("pellets": the collision vertices of every body, no volume maps; only
 * the ISPH solver steps with them, as in the reference). */
#define ALV_VOLUME_MAP 0
#define ALV_PELLETS 1

/* Pile::Pile (pile.hpp:151-192); ALV_PELLETS with max_num_pellets = 10000,
 * the reference's default. */
int alv_pile_create(alv_pile** out, unsigned max_num_contacts,
                    int volume_method);
/* Pile::Pile with max_num_pellets (pile.hpp:151-166) */
int alv_pile_create_ex(alv_pile** out, unsigned max_num_contacts,
                       int volume_method, unsigned max_num_pellets);
/* num_pellets_ and max_num_pellets_ (pile.hpp:124-130) */
int alv_pile_num_pellets(const alv_pile* pile, unsigned* num_pellets,
                         unsigned* max_num_pellets);
/* Pile::hint_identical_sequence (pile.hpp:637-639): bodies [begin, end)
 * share one shape; the pellet contact search tests them in one run. */
int alv_pile_hint_identical_sequence(alv_pile* pile, unsigned begin,
                                     unsigned end);
int alv_pile_destroy(alv_pile* pile);
/* Pile::add (pile.hpp:217-352): computes the map domain and resolution
 * (calculate_domain_and_resolution, pile.hpp:40-65) and, for volume maps,
 * builds the map on the device (update_volume_field, runner.hpp:1660-1694;
 * build_grid, pile.hpp:493-623).  A pellet pile builds no map. */
int alv_pile_add(alv_pile* pile, const alv_const* cn,
                 const alv_distance* distance, const unsigned resolution[3],
                 float cell_width, float margin, float sign, float mass,
                 float restitution, float friction, const float inertia[3],
                 const float x[3], const float q[4], unsigned* boundary_id);
int alv_pile_get_size(const alv_pile* pile, unsigned* size);
/* Pile::replace (pile.hpp:423-492): boundary b gets a new distance, map and
 * mass properties, x and q, with v = omega = 0, as alv_pile_add would build
 * it; solvers use it from their next step.  (Collision vertices are set
 * separately, alv_pile_set_collision_vertices.) */
int alv_pile_replace(alv_pile* pile, unsigned b, const alv_const* cn,
                     const alv_distance* distance, const unsigned resolution[3],
                     float cell_width, float margin, float sign, float mass,
                     float restitution, float friction, const float inertia[3],
                     const float x[3], const float q[4]);
/* mass_, restitution_, friction_, inertia_tensor_ of body b (pile.hpp:96-102,
 * read-write in the reference's bindings, pyalluvion.cu:466-473) */
int alv_pile_set_body(alv_pile* pile, unsigned b, float mass, float restitution,
                      float friction, const float inertia[3]);
int alv_pile_get_body(const alv_pile* pile, unsigned b, float* mass, float* restitution,
                      float* friction, float inertia[3]);
/* Pile::calculate_cfl_v2 (pile.hpp:816-825): the largest
 * |omega x (max_distance, 0, 0) + v|^2 over the bodies (lowest float with
 * none). */
int alv_pile_calculate_cfl_v2(const alv_pile* pile, float* max_v2);
/* Pile::compute_mask (pile.hpp:1048-1065, analytic distances,
 * compute_boundary_mask runner.hpp:4385-4397): for the device arrays
 * sample_x [n] float3 and mask [n] float, mask[i] *= (sign * d >= threshold)
 * with d the signed distance of sample i in body b's frame (its current
 * host kinematics). */
int alv_pile_compute_mask(alv_pile* pile, unsigned b, float distance_threshold,
                          const float* sample_x, float* mask, unsigned num_samples);
/* Host kinematics x_, v_, q_, omega_ (pile.hpp:103-110) */
int alv_pile_set_kinematics(alv_pile* pile, unsigned b, const float x[3],
                            const float v[3], const float q[4],
                            const float omega[3]);
int alv_pile_get_kinematics(const alv_pile* pile, unsigned b, float x[3],
                            float v[3], float q[4], float omega[3],
                            float force[3], float torque[3]);
int alv_pile_set_gravity(alv_pile* pile, const float g[3]);
/* Pile::integrate_kinematics (pile.hpp:794-815) */
int alv_pile_integrate_kinematics(alv_pile* pile, float dt);

/* Device copies of the bodies' kinematics, refreshed at the start of every
 * solver step (Pile::copy_kinematics_to_device, pile.hpp:751-756): x, v,
 * omega are [B][3], q is [B][4] (x, y, z, w).  The pointers stay valid
 * while the pile lives. */
int alv_pile_copy_kinematics_to_device(alv_pile* pile);
int alv_pile_kinematics_device(alv_pile* pile, float** x, float** v, float** q, float** omega);

/* ---- rigid contacts (contact.hpp:7-24, pile.hpp:826-979) ------------------ */
/* Contact record; the layout of the reference's Contact<float3, float>. */
typedef struct alv_contact {
  unsigned i, j;
  float cp_i[3], cp_j[3];
  float n[3], t[3];
  float iiwi_diag[3], iiwi_off_diag[3];
  float iiwj_diag[3], iiwj_off_diag[3];
  float friction, nkninv, pmax, goalu, impulse_sum;
} alv_contact;
/* Collision vertices of body b in its own frame (the vertices of the
 * reference's collision_mesh argument of Pile::add, pile.hpp:322-331, or
 * its pellets, pile.hpp:259-276): `xyz` is a host array of 3*n floats;
 * n = 0 removes them.  A pellet pile fails with ALV_ERR_ARG when the pellet
 * count would exceed max_num_pellets. */
int alv_pile_set_collision_vertices(alv_pile* pile, unsigned b, const float* xyz,
                                    unsigned n);
/* Pile::find_contacts (pile.hpp:826-904).  i < 0: every pair (i, j), the
 * overload the DFSPH step calls; i >= 0, j < 0: body i against all others;
 * both >= 0: that pair.  Pairs with i == j or two bodies of zero mass are
 * skipped, as are pairs whose body j is a box shell (the reference's
 * collision_test has no box-shell variant, runner.hpp:5891-5977).
 * `contact_tolerance` is Const::contact_tolerance.  Contacts are stored in
 * (i, j, vertex) order; *num_contacts is the number found, which may exceed
 * max_num_contacts (only that many are kept). */
int alv_pile_find_contacts(alv_pile* pile, int i, int j, float contact_tolerance,
                           unsigned* num_contacts);
/* Pile::find_contacts(particle_x, max_num_beads), pellet branch
 * (pile.hpp:858-886, collision_test_with_pellets runner.hpp:3627-3727):
 * the world pellets, rows max_num_beads .. of the device array particle_x
 * (float3, the solver's position field), against every body (or hinted run
 * of identical bodies) with the kinematics of the last device copy; pairs
 * i == j and of two massless bodies are skipped and box shells have no
 * test.  Contacts in (run, pellet, j) order.  A volume-map pile runs the
 * all-pairs search instead, as the reference does. */
int alv_pile_find_contacts_pellets(alv_pile* pile, const float* particle_x,
                                   unsigned max_num_beads, float contact_tolerance,
                                   unsigned* num_contacts);
/* Pile::solve_contacts (pile.hpp:906-979): five sequential impulse sweeps
 * over the stored contacts, changing the host v and omega of the bodies. */
int alv_pile_solve_contacts(alv_pile* pile);
/* Copies up to `capacity` stored contacts (as left by the last find or
 * solve, impulse sums included) to host memory; *n = the number copied. */
int alv_pile_get_contacts(const alv_pile* pile, alv_contact* out, unsigned capacity,
                          unsigned* n);
/* Volume map of boundary b (device pointer) and its grid geometry. */
int alv_pile_get_volume_grid(const alv_pile* pile, unsigned b, float** nodes,
                             unsigned* num_nodes);
/* Distance map of boundary b (Pile::get_distance_grids, pile.hpp:981-988):
 * device pointer and node count; null and 0 for an analytic distance, whose
 * reference distance grid is empty. */
int alv_pile_get_distance_grid(const alv_pile* pile, unsigned b, float** nodes,
                               unsigned* num_nodes);
int alv_pile_get_grid_info(const alv_pile* pile, unsigned b,
                           float domain_min[3], float domain_max[3],
                           unsigned resolution[3], float cell_size[3],
                           float* sign, float* mass);

/* ---- solvers: SolverDf<float> / SolverIi<float> --------------------------- */
typedef struct alv_solver alv_solver;
#define ALV_SOLVER_DF 0
#define ALV_SOLVER_II 1
#define ALV_SOLVER_I 2 /* SolverI: pressure-Laplacian ISPH (solver_i.hpp:8-353) */

/* Public fields of Solver / SolverDf / SolverIi (solver.hpp:261-272,
 * solver_df.hpp:372-379, solver_ii.hpp:259-262). */
typedef struct alv_solver_settings {
  unsigned num_particles;
  float t;
  float dt;
  float initial_dt;
  float max_dt;
  float min_dt;
  float cfl;
  float density_change_tolerance;
  float density_error_tolerance;
  unsigned min_density_solve;
  unsigned max_density_solve;
  unsigned min_divergence_solve;
  unsigned max_divergence_solve;
  int enable_divergence_solve;
  int enable_density_solve;
} alv_solver_settings;

/* Report fields (solver.hpp:274-279, solver_df.hpp:381-385).  A DFSPH step
 * with a solve disabled leaves that solve's count and mean as they were, as
 * the reference sets them inside the solve (solver_df.hpp:129-130, 294-295). */
typedef struct alv_solver_report {
  unsigned num_divergence_solve;
  unsigned num_density_solve;
  float mean_density_change;
  float mean_density_error;
  float particle_max_v2;
  float pile_max_v2;
  float max_v2;
  float cfl_dt;
  float utilized_cfl;
  float t;
  float dt;
} alv_solver_report;

/* SolverDf / SolverIi / SolverI constructors (solver.hpp:19-94,
 * solver_df.hpp:48-67, solver_ii.hpp:40-65, solver_i.hpp:48-66).  The pile
 * must outlive the solver. */
int alv_solver_create(alv_solver** out, int kind, alv_pile* pile,
                      unsigned max_num_particles, const alv_const* cn,
                      const alv_consti* cni);
int alv_solver_destroy(alv_solver* solver);

/* Batch of independent DFSPH or IISPH scenes in one solver.  No counterpart
 * in the reference, which holds one scene per process (its constants are
 * process-global, constants.hpp:141-143).  Every scene has its own
 * particles, neighbour grid and pile of static boundaries; the constants, the
 * neighbour caps of `cni` and the settings are shared.  One launch of each
 * kernel advances every scene, and every scene keeps its own time step,
 * convergence test and iteration count: a converged scene idles while the
 * others iterate.  Particle arrays hold the scenes one after another; scene k
 * occupies [first, first + count) (alv_solver_scene_range) of every field. */
typedef struct {
  alv_pile* pile;
  unsigned num_particles;
  unsigned grid_res[3];
  int grid_offset[3];
} alv_scene_desc;
int alv_solver_create_batch(alv_solver** out, int kind, unsigned num_scenes,
                            const alv_scene_desc* scenes, const alv_const* cn,
                            const alv_consti* cni);
int alv_solver_num_scenes(const alv_solver* solver, unsigned* num_scenes);
int alv_solver_scene_range(const alv_solver* solver, unsigned scene, unsigned* first,
                           unsigned* count);
/* Report of one scene after the last step (t and dt are the scene's own). */
int alv_solver_scene_report(const alv_solver* solver, unsigned scene,
                            alv_solver_report* report);

/* Store::copy_cn_external (store.hpp:142-164): constants used from the next
 * step on. */
int alv_solver_set_const(alv_solver* solver, const alv_const* cn,
                         const alv_consti* cni);
int alv_solver_set_settings(alv_solver* solver,
                            const alv_solver_settings* settings);
int alv_solver_get_settings(const alv_solver* solver,
                            alv_solver_settings* settings);
int alv_solver_get_report(const alv_solver* solver,
                          alv_solver_report* report);

/* Particle arrays (solver.hpp:286-305, solver_df.hpp:367-370,
 * solver_ii.hpp:250-257) in user (particle id) order.  *dptr is a device
 * pointer valid for the solver's lifetime; n_elements and n_components give
 * its shape ([B, N] arrays report B*N elements).  Arrays that the step
 * computes in cell-sorted order are brought to user order on request. */
#define ALV_FIELD_X 0
#define ALV_FIELD_V 1
#define ALV_FIELD_A 2
#define ALV_FIELD_DENSITY 3
#define ALV_FIELD_DFSPH_FACTOR 4
#define ALV_FIELD_KAPPA 5
#define ALV_FIELD_KAPPA_V 6
#define ALV_FIELD_DENSITY_ADV 7
#define ALV_FIELD_NUM_NEIGHBORS 8
#define ALV_FIELD_PRESSURE 9
#define ALV_FIELD_PID_LENGTH 10
#define ALV_FIELD_GRID_ANOMALY 11
#define ALV_FIELD_CELL_START 12
/* Per-particle u32 bit mask of the boundaries with non-zero volume
 * (bit b = boundary b), in cell-sorted order; for statistics. */
#define ALV_FIELD_BOUNDARY_MASK 13
/* Reference-layout arrays made on request from the last step (or neighbour
 * search), in user order; the buffers are allocated on first request and
 * belong to the solver.
 * particle_neighbors (solver.hpp:78-85, filled by make_neighbor_list,
 * runner.hpp:1896-1899): [max_n][K] float4 (x_i - x_j, bits of j), slots
 * past the count zero; */
#define ALV_FIELD_NEIGHBORS 14
/* particle_boundary, particle_boundary_kernel (solver.hpp:39-42,
 * runner.hpp:1982-2025): [B][max_n] float4 (x_b, Psi) and (Psi grad W,
 * Psi W), zero where the particle does not see boundary b; */
#define ALV_FIELD_BOUNDARY 15
#define ALV_FIELD_BOUNDARY_KERNEL 16
/* particle_force, particle_torque (solver.hpp:43-46): [B][max_n] float3,
 * the step's per-particle forces on each boundary; rows of static
 * boundaries are zero unless alv_solver_set_static_boundary_forces is on; */
#define ALV_FIELD_FORCE 17
#define ALV_FIELD_TORQUE 18
/* SolverIi's particle_last_pressure, particle_aii, particle_adv_density,
 * particle_density_err ([max_n] float) and particle_dii, particle_dij_pj,
 * particle_pressure_accel ([max_n] float3) (solver_ii.hpp:50-64). */
#define ALV_FIELD_LAST_PRESSURE 19
#define ALV_FIELD_AII 20
#define ALV_FIELD_ADV_DENSITY 21
#define ALV_FIELD_DENSITY_ERR 22
#define ALV_FIELD_DII 23
#define ALV_FIELD_DIJ_PJ 24
#define ALV_FIELD_PRESSURE_ACCEL 25
int alv_solver_get_field(alv_solver* solver, int field, void** dptr,
                         size_t* n_elements, unsigned* n_components);

/* SolverDf::step<wrap> (solver_df.hpp:74-359) / SolverIi::step<wrap>
 * (solver_ii.hpp:76-243).  Periodicity follows alv_const.wrap_length; the
 * `wrap` argument of the reference (y only) is implied by it. */
int alv_solver_step(alv_solver* solver);
/* Step execution (no counterpart in the reference): with `enable` the DFSPH
 * step runs as three HIP graphs (neighbour search and divergence warm start,
 * viscosity to density warm start, final integration), captured on first use
 * and replayed; each graph holds the last step's Jacobi iteration count plus
 * `slack` iterations, iterations past convergence return at once, and the
 * host adds more only when a scene has not converged.  Results are identical
 * with and without graphs.  Per-kernel profiling uses direct launches. */
int alv_solver_set_graph(alv_solver* solver, int enable, unsigned slack);
/* Neighbour passes that copy each 256-particle block's neighbourhood (three
 * contiguous ranges of sorted particles) into LDS before their loops (no
 * counterpart in the reference; on by default).  Blocks whose neighbourhood
 * does not fit, and periodic grids, read neighbours from global memory.
 * Results are identical either way. */
int alv_solver_set_tiles(alv_solver* solver, int enable);
/* The reference accumulates per-particle boundary forces and torques for
 * every boundary in every pressure and viscosity pass (runner.hpp:3252-3363,
 * 2150-2249) but only moving bodies use them; here only bodies with mass
 * accumulate by default.  With `enable`, static boundaries accumulate too
 * from the next step on (ALV_FIELD_FORCE / ALV_FIELD_TORQUE rows). */
int alv_solver_set_static_boundary_forces(alv_solver* solver, int enable);
/* Solver::compute_all_boundaries (solver.hpp:170-185) followed by
 * Solver::update_particle_neighbors<wrap> (solver.hpp:225-260). */
int alv_solver_update_particle_neighbors(alv_solver* solver);
/* The density-change sum of compute_velocity_of_density_change
 * (runner.hpp:3297-3314, via compute_density_change 3192-3228) for every
 * particle, into particle_density_adv, without the dfsph_factor update.
 * Uses the neighbour list and boundary records of the last
 * update_particle_neighbors. */
int alv_solver_compute_density_change(alv_solver* solver);
/* integrate_velocity<wrap> (runner.hpp:3613-3623): x += dt v with the
 * solver's dt, wrapped on every periodic axis. */
int alv_solver_integrate_velocity(alv_solver* solver);
/* Solver::reset_solving_var (solver_df.hpp:361-365, solver_ii.hpp:245-248)
 * and Solver::reset_t (solver.hpp:148-152). */
int alv_solver_reset_solving_var(alv_solver* solver);
int alv_solver_reset_t(alv_solver* solver);
/* Neighbour list in the reference layout (particle_neighbors /
 * particle_num_neighbors, solver.hpp:78-85): host buffers ids[N*Kmax]
 * (neighbour particle ids, canonical order) and counts[N]. */
int alv_solver_get_neighbors(alv_solver* solver, unsigned* ids,
                             unsigned* counts);
/* Pellet lists of a pellet solver in the reference layout
 * (particle_boundary_neighbors / particle_num_boundary_neighbors and the
 * pellets' rows of particle_neighbors, solver.hpp:78-85): which = 0, the
 * pellets near each fluid particle, ids[max_n*Kmax] as particle ids
 * max_n + k, counts[max_n]; which = 1, the fluid particles near each pellet,
 * ids[max_num_pellets*Kmax] as fluid ids, counts[max_num_pellets]. */
int alv_solver_get_pellet_neighbors(alv_solver* solver, int which,
                                    unsigned* ids, unsigned* counts);
/* Solver::transform_all_pellets (solver.hpp:204-223): places the pile's
 * pellets into rows max_n .. of particle_x / particle_v with the bodies'
 * current host kinematics (a no-op without pellets).  The ISPH step does it
 * at its start; update_particle_neighbors uses the rows as they are. */
int alv_solver_transform_all_pellets(alv_solver* solver);
/* Cell counts (pid_length) [C], cell offsets [C+1] and grid_anomaly[3] into
 * host buffers. */
int alv_solver_get_grid(alv_solver* solver, unsigned* cell_count,
                        unsigned* cell_start, unsigned* anomaly);

/* ---- per-kernel timing (Runner::launch statistics, runner.hpp:4835-4859) -- */
/* HIP events around every kernel launch of this process; off by default.
 * Times are gathered at the end of each step, so enabling it adds no host
 * synchronisation inside a step. */
int alv_profiling_enable(int enable);
int alv_profiling_reset(void);
int alv_profiling_num_kernels(unsigned* n);
int alv_profiling_kernel(unsigned index, const char** name, double* total_ms,
                         unsigned long long* launches);

/* ---- device arithmetic checks (no counterpart in the reference) ---------- */
/* The kernels' correctly rounded arithmetic, element-wise on the GPU over
 * host buffers a[n] (and b[n]), for callers that check it against IEEE
 * results: op 0 sqrt(a), op 1 a / b by the known-divisor division, op 2 the
 * cubic-spline gradient scale of the displacement (a, 0, 0), op 3 the cubic
 * spline W of the squared distance a, with the kernel radius of cn; op 4
 * 1 / a, op 5 a / b by the general division (the reciprocal refined to the
 * correctly rounded one); op 6 the gradient scale of op 2 with the
 * compiler's IEEE division in its outer branch. */
int alv_check_math(const alv_const* cn, int op, const float* a, const float* b, float* out,
                   size_t n);

/* ---- stand-alone volume-map kernels (tests/field_op.cu, field_gen.cu) ----- */
/* index_to_node_position over all nodes of a grid (runner.hpp:1131-1178):
 * out[3*num_nodes] host buffer. */
int alv_node_positions(const float domain_min[3], const unsigned resolution[3],
                       const float cell_size[3], float* out,
                       unsigned* num_nodes);
/* update_volume_field, distance-map variant (runner.hpp:1616-1658), on host
 * buffers: distance[num_nodes] in, volume[num_nodes] out. */
int alv_volume_from_distance_map(const alv_const* cn,
                                 const float domain_min[3],
                                 const float domain_max[3],
                                 const unsigned resolution[3],
                                 const float cell_size[3], float sign,
                                 const float* distance, float* volume,
                                 unsigned num_nodes);

/* ---- learned warm start --------------------------------------------------
 * No counterpart in the reference.  A small network (8 -> 64 -> 64 -> 64 ->
 * 2, SiLU; bfloat16 weights and activations, FP32 accumulation on matrix
 * cores) predicts each particle's converged divergence and density stiffness
 * (the stored forms kappa_v and kappa of runner.hpp:3417-3418,3606-3611) from
 * eight local features: rho/rho0 - 1, the density change rho_dot of
 * compute_density_change (runner.hpp:3192-3228), the neighbour count, the
 * summed boundary volume sum_b Psi_b, the previous kappa_v and kappa, |v| and
 * dt.  The solver uses the predictions in place of the warm starts of
 * warm_start_divergence_solve_0 and warm_start_pressure_solve0
 * (runner.hpp:3230-3250,3430-3450); the convergence tests are unchanged.
 *
 * Parameters are a flat FP32 vector W1[64][8] b1 W2[64][64] b2 W3[64][64] b3
 * W4[2][64] b4 (9026 values).  Normalisation: 20 floats, in_shift[8],
 * in_scale[8], out_shift[2], out_scale[2]; the network sees
 * (f - in_shift) * in_scale and its outputs y map to y * out_scale +
 * out_shift.  Device pointers are marked dev_. */
typedef struct alv_mlp alv_mlp;
int alv_mlp_num_params(void);
/* Glorot-uniform weights from SplitMix64(seed), zero biases, identity
 * normalisation. */
int alv_mlp_create(alv_mlp** out, unsigned long long seed);
int alv_mlp_destroy(alv_mlp* net);
int alv_mlp_set_params(alv_mlp* net, const float* host_params);
int alv_mlp_get_params(alv_mlp* net, float* host_params);
int alv_mlp_set_normalization(alv_mlp* net, const float* norm20);
int alv_mlp_get_normalization(alv_mlp* net, float* norm20);
/* mean / standard deviation of the inputs and targets of a data set
 * (dev_x [n][8], dev_y [n][2]) over a strided sample of at most 2^20 pairs */
int alv_mlp_fit_normalization(alv_mlp* net, const float* dev_x, const float* dev_y, size_t n);
/* raw features dev_x [n][8] -> de-normalised predictions dev_y [n][2] */
int alv_mlp_forward(alv_mlp* net, const float* dev_x, size_t n, float* dev_y);
/* mean squared error of normalised batch (dev_xn [n][8], dev_tn [n][2]) and
 * its gradient (host_grad[9026]); n a positive multiple of 2048 */
int alv_mlp_loss_grad(alv_mlp* net, const float* dev_xn, const float* dev_tn, size_t n,
                      double* loss, float* host_grad);
/* `steps` Adam steps (learning rate lr) on random mini-batches of `batch`
 * pairs (a multiple of 2048) drawn from the data set; host_loss[steps]
 * (optional) receives the batch losses */
int alv_mlp_train(alv_mlp* net, const float* dev_x, const float* dev_y, size_t n,
                  unsigned batch, unsigned steps, float lr, unsigned long long seed,
                  float* host_loss);
/* Use a network for the DFSPH warm starts of later steps: the divergence
 * warm start (use_divergence), the density warm start (use_density), or
 * both.  A null network or two zero flags restore the reference's warm
 * starts.  The solver does not own the network. */
int alv_solver_set_warm_start_net(alv_solver* solver, alv_mlp* net, int use_divergence,
                                  int use_density);
/* Record the eight warm-start features during later steps. */
int alv_solver_set_feature_log(alv_solver* solver, int enable);
/* Training pairs of the last step (num_particles of them, in an internal
 * particle order): features dev_x [n][8] and the converged stored
 * (kappa_v, kappa) dev_y [n][2]. */
int alv_solver_export_training_pairs(alv_solver* solver, float* dev_x, float* dev_y);

/* ---- sampling and statistics --------------------------------------------- */
/* All pointers are device pointers; launches are on the default stream.
 * Sample neighbour lists use the reference's format: [num_samples][K]
 * float4, K = max_num_neighbors_per_particle, entry = (x_sample - x_j,
 * particle id j in the bits of w).
 *
 * make_neighbor_list for sample points (runner.hpp:1869-1911, launched by
 * Runner::launch_make_neighbor_list, runner.hpp:5604-5620): against the grid
 * and particle positions of the solver's last neighbour search (a step or
 * update_particle_neighbors must have run).  A particle whose id equals the
 * sample index is skipped, as in the reference.  A pellet solver's pellets
 * are in the grid too (ids max_n + k, after the cell's fluid particles).
 * Sets grid_anomaly[2] when a sample has more than K neighbours (the last
 * slot keeps the latest).
 * max_neighbors (the row length of sample_neighbors) must equal the solver's
 * max_num_neighbors_per_particle. */
int alv_solver_sample_neighbors(alv_solver* solver, const float* sample_x,
                                unsigned num_samples, unsigned max_neighbors,
                                float* sample_neighbors, unsigned* sample_num_neighbors);
/* make_bead_pellet_neighbor_list<wrap> (runner.hpp:1913-1971; Runner
 * launcher runner.hpp:5622-5641) against the solver's grid of the last
 * neighbour search: samples offset .. offset+num_samples-1 of the device
 * array sample_x (float3), their neighbours within the kernel radius split
 * by id into beads (id < max_num_beads) and pellets (ids max_n + k), in the
 * reference order, skipping the id equal to the sample's index; rows
 * (float4 xixj + id bits, max_neighbors each) and counts are indexed by the
 * sample's index (offset included). */
int alv_solver_sample_bead_pellet_neighbors(alv_solver* solver, const float* sample_x,
                                            unsigned num_samples, unsigned offset,
                                            unsigned max_num_beads, unsigned max_neighbors,
                                            float* bead_neighbors, unsigned* bead_counts,
                                            float* pellet_neighbors, unsigned* pellet_counts);
/* Sums over reference-layout lists (device float4 rows of K entries):
 * compute_particle_boundary_with_pellets (runner.hpp:2027-2050):
 * boundary_kernel[i] = vol (sum grad W, sum W) over the pellet list;
 * compute_density_with_pellets (runner.hpp:2074-2098): m (W(0) + sum W +
 * sum W) over both lists; compute_dfsph_factor_with_pellets
 * (runner.hpp:3156-3190): the DFSPH factor, the pellet entries' vectors
 * taken as -(grad p_j) as the reference reads them. */
/* Sampling with pellet lists: what 0 = sample_position_density_with_pellets
 * (runner.hpp:2122-2147, out[ns] = m (sum W + sum W)), 1 =
 * compute_sample_velocity_with_pellets (runner.hpp:4242-4277), 2 =
 * compute_sample_vorticity_with_pellets (runner.hpp:4325-4359; sample_v the
 * samples' velocity); out is float3 [ns] for 1 and 2.  Fluid entries are
 * weighted m / rho_j, pellet entries by the particle volume; particle_v
 * holds the pellets' velocities at rows max_n + k (the solver's field). */
int alv_sample_with_pellets(const alv_const* cn, int what, const float* neighbors,
                            const unsigned* counts, const float* pellet_neighbors,
                            const unsigned* pellet_counts, unsigned K, unsigned num_samples,
                            const float* particle_density, const float* particle_v,
                            const float* sample_v, float* out);
int alv_particle_boundary_with_pellets(const alv_const* cn, const float* pellet_neighbors,
                                       const unsigned* pellet_counts, unsigned K, unsigned n,
                                       float* boundary_kernel);
int alv_density_with_pellets(const alv_const* cn, const float* neighbors, const unsigned* counts,
                             const float* pellet_neighbors, const unsigned* pellet_counts,
                             unsigned K, unsigned n, float* density);
int alv_dfsph_factor_with_pellets(const alv_const* cn, const float* neighbors,
                                  const unsigned* counts, const float* pellet_neighbors,
                                  const unsigned* pellet_counts, unsigned K, unsigned n,
                                  float* factor);
/* Solver::sample_all_boundaries (solver.hpp:186-202): boundary records of the
 * sample points ([B][num_samples] float4 each) with the pile's current
 * kinematics. */
int alv_solver_sample_all_boundaries(alv_solver* solver, const float* sample_x,
                                     unsigned num_samples, float* sample_boundary,
                                     float* sample_boundary_kernel);
/* sample_fluid (runner.hpp:4173-4192): sum_j m / rho_j W(x_sample - x_j) q_j
 * for a scalar (components = 1) or 3-vector quantity in particle order. */
int alv_sample_fluid(const alv_const* cn, const float* sample_neighbors,
                     const unsigned* sample_num_neighbors, unsigned max_neighbors,
                     unsigned num_samples, const float* particle_density,
                     const float* particle_quantity, unsigned components,
                     float* sample_quantity);
/* sample_fluid_density (runner.hpp:4193-4209): rho0 sum_j V W */
int alv_sample_fluid_density(const alv_const* cn, const float* sample_neighbors,
                             const unsigned* sample_num_neighbors, unsigned max_neighbors,
                             unsigned num_samples, float* sample_density);
/* sample_position_density (runner.hpp:2100-2120): as sample_fluid_density
 * plus the boundary terms sum_b Psi_b W_b of the sample boundary kernels
 * ([num_boundaries][num_samples] float4). */
int alv_sample_position_density(const alv_const* cn, const float* sample_neighbors,
                                const unsigned* sample_num_neighbors, unsigned max_neighbors,
                                unsigned num_samples, const float* sample_boundary_kernel,
                                unsigned num_boundaries, float* sample_density);
/* compute_sample_velocity (runner.hpp:4211-4241) and
 * compute_sample_vorticity (runner.hpp:4289-4324); boundary terms use the
 * sample boundary records and the bodies' x, v, omega ([B][3], e.g. the
 * pile's device kinematics). */
int alv_sample_velocity(const alv_const* cn, const float* sample_neighbors,
                        const unsigned* sample_num_neighbors, unsigned max_neighbors,
                        unsigned num_samples, const float* particle_density,
                        const float* particle_v, const float* sample_boundary,
                        const float* sample_boundary_kernel, unsigned num_boundaries,
                        const float* rigid_x, const float* rigid_v, const float* rigid_omega,
                        float* sample_v);
int alv_sample_vorticity(const alv_const* cn, const float* sample_neighbors,
                         const unsigned* sample_num_neighbors, unsigned max_neighbors,
                         unsigned num_samples, const float* particle_density,
                         const float* particle_v, const float* sample_v,
                         const float* sample_boundary, const float* sample_boundary_kernel,
                         unsigned num_boundaries, const float* rigid_x, const float* rigid_v,
                         const float* rigid_omega, float* sample_vorticity);
/* Runner::launch_histogram256[_with_mask] (runner.hpp:5849-5890): 256-bin
 * histogram of floor((v - v_min) * 256 / (v_max - v_min)) mod 256 (negative
 * arguments in bin 0), counting element i only where mask[i] == 1 when a mask
 * is given; histogram is [256] u32. */
int alv_histogram256(const float* v, const float* mask, unsigned n, float v_min, float v_max,
                     unsigned* histogram);
/* Runner::calculate_kl_divergence (runner.hpp:4803-4815, KLDivergence
 * :746-762) of two histograms; the result is written to *result (host). */
int alv_kl_divergence(const unsigned* histogram_p, const unsigned* histogram_q, unsigned n_p,
                      unsigned n_q, float q_lower_bound, unsigned num_bins, float* result);

#ifdef __cplusplus
}
#endif
#endif
