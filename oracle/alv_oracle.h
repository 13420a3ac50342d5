/*
 * alv_oracle.h -- C interface of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a plain C++17/OpenMP restatement
 * of the reference's SPH solver-core algorithm (kennychufk/alluvion,
 * include/alluvion/runner.hpp, solver.hpp, solver_df.hpp, solver_ii.hpp,
 * pile.hpp).  It exists to check the HIP product path and to time the CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; the product library never links or calls it.
 *
 * Every per-particle loop is independent, so results do not depend on the
 * OpenMP thread count.  Reductions are serial FP64 sums in particle order.
 */
#ifndef ALV_ORACLE_H
#define ALV_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Physical and grid constants (mirror of Const<TF> + Consti,
 * reference constants.hpp:40-137).  Derived constants (h^2, cubic k, l, W(0),
 * particle volume, wrap bounds) are computed inside the oracle in the working
 * precision, exactly as the reference setters do. */
typedef struct {
  double kernel_radius;
  double particle_radius;
  double particle_mass;
  double density0;
  double viscosity;
  double boundary_viscosity;
  double dfsph_factor_epsilon;
  double gravity[3];
  /* Periodic wrap per axis (0 = off).  The reference wraps y only
   * (runner.hpp:689-707); x and z are the extension used by the
   * uniform-random periodic stress scene. */
  double wrap_length[3];
  int grid_res[3];
  int grid_offset[3];
  unsigned max_num_neighbors_per_particle;
  unsigned max_num_particles_per_cell;
} orc_params;

/* Solver controls (mirror of SolverDf / SolverIi / Solver public fields,
 * solver_df.hpp:62-67, solver.hpp:261-272). */
typedef struct {
  double dt;
  double max_dt;
  double min_dt;
  double cfl;
  double density_change_tolerance;
  double density_error_tolerance;
  unsigned min_density_solve;
  unsigned max_density_solve;
  unsigned min_divergence_solve;
  unsigned max_divergence_solve;
  int enable_divergence_solve;
  int enable_density_solve;
} orc_settings;

typedef struct {
  unsigned num_divergence_solve;
  unsigned num_density_solve;
  double mean_density_change;
  double mean_density_error;
  double particle_max_v2;
  double pile_max_v2;
  double max_v2;
  double cfl_dt;
  double utilized_cfl;
  double t;
  double dt;
} orc_report;

/* Analytic distance types (reference dg/box_distance.hpp,
 * dg/box_shell_distance.hpp, dg/sphere_distance.hpp), and a distance sampled
 * on the boundary's map nodes (ORC_MAP: the reference's grid path for
 * MeshDistance, runner.hpp:1696-1736,1773-1825,2005-2025; the node values
 * are set with orc_set_distance_map). */
enum { ORC_BOX = 0, ORC_BOX_SHELL = 1, ORC_SPHERE = 2, ORC_MAP = 3 };

typedef struct {
  int type;
  /* box: widths[3], outset; shell: inner widths[3], thickness, outset;
   * sphere: radius */
  double p[5];
  double sign;
  unsigned resolution[3];
  double margin; /* negative = reference default (h if sign<0 else 2h) */
  double mass;
  double inertia[3];
  double x[3];
  double q[4]; /* (x, y, z, w) */
  double aabb_min[3], aabb_max[3]; /* ORC_MAP: the distance's bounding box */
} orc_boundary_desc;

/* kind: 0 = DFSPH (SolverDf), 1 = IISPH (SolverIi), 2 = ISPH (SolverI).
 * precision: 4 or 8. */
void* orc_create(int precision, int kind, unsigned max_num_particles,
                 const orc_params* params);
void orc_destroy(void* h);
int orc_set_settings(void* h, const orc_settings* s);
int orc_set_num_particles(void* h, unsigned n);

/* Adds an analytic boundary; computes domain/resolution/cell size like
 * Pile::add (pile.hpp:40-78,285-352).  The volume map starts empty. */
int orc_add_boundary(void* h, const orc_boundary_desc* d);
unsigned orc_num_nodes(void* h, unsigned b);
int orc_get_grid_info(void* h, unsigned b, double* domain_min,
                      double* domain_max, unsigned* resolution,
                      double* cell_size);
/* Volume map: build with the Peirce rule (runner.hpp:1660-1694) over nodes
 * [begin, end), or set / get the node values (working precision). */
int orc_build_volume_map(void* h, unsigned b, unsigned begin, unsigned end);
int orc_set_volume_map(void* h, unsigned b, const void* nodes);
int orc_get_volume_map(void* h, unsigned b, void* out);
/* Distance map of an ORC_MAP boundary (working precision, one value per map
 * node); orc_build_volume_map then builds the volume map from it. */
int orc_set_distance_map(void* h, unsigned b, const void* nodes);
/* interpolate_and_derive (runner.hpp:1773-1791) on a stand-alone map:
 * value and gradient at each of n points (nodes in double). */
int orc_map_value_gradient(const double* domain_min, const double* domain_max,
                           const unsigned* res, const double* cell_size,
                           const double* nodes, size_t n, const double* points,
                           double* value, double* gradient, int* inside);
int orc_set_rigid(void* h, unsigned b, const double* x, const double* v,
                  const double* q, const double* omega);
int orc_get_rigid(void* h, unsigned b, double* x, double* v, double* q,
                  double* omega, double* force, double* torque);
int orc_set_gravity_pile(void* h, const double* g);

/* Rigid contacts (contact.hpp:7-24, runner.hpp:3833-3914 analytic variant,
 * pile.hpp:826-979).  Collision vertices of body b in its frame (3*n), with
 * the body's restitution and friction; Const::contact_tolerance and
 * Consti::max_num_contacts.  Contacts are found in (i, j, vertex) order:
 * i < 0 all pairs, j < 0 body i against the others.  A contact record has
 * ORC_CONTACT_FIELDS doubles: i, j, cp_i, cp_j, n, t, iiwi_diag,
 * iiwi_off_diag, iiwj_diag, iiwj_off_diag, friction, nkninv, pmax, goalu,
 * impulse_sum.  The step finds and solves contacts after integrating the
 * bodies whenever any body has collision vertices. */
/* Sampling against the grid of the last neighbour search (make_neighbor_list
 * on sample points, runner.hpp:1869-1911; sample kernels runner.hpp:2100-2120,
 * 4173-4324).  ids and xixj are [ns][Kmax] (xixj packed xyz), counts [ns];
 * boundary records [B][ns][4]; quantity [n][dim] in particle order.
 * orc_sample_velocity computes the sample vorticity when sample_v is given. */
int orc_sample_neighbors(void* h, const double* sample_x, unsigned ns, unsigned* ids,
                         double* xixj, unsigned* counts);
int orc_sample_boundary(void* h, const double* sample_x, unsigned ns, double* rec, double* ker);
int orc_sample_fluid(void* h, const unsigned* ids, const double* xixj, const unsigned* counts,
                     unsigned ns, const double* quantity, unsigned dim, double* out);
int orc_sample_density(void* h, const double* xixj, const unsigned* counts, unsigned ns,
                       const double* ker, unsigned nb, double* out);
int orc_sample_velocity(void* h, const unsigned* ids, const double* xixj, const unsigned* counts,
                        unsigned ns, const double* sample_v, const double* rec, const double* ker,
                        unsigned nb, const double* rigid_x, const double* rigid_v,
                        const double* rigid_omega, double* out);

#define ORC_CONTACT_FIELDS 31
int orc_set_collision(void* h, unsigned b, double restitution, double friction,
                      const double* xyz, unsigned n);
int orc_set_contact_params(void* h, double contact_tolerance, unsigned max_num_contacts);
int orc_find_contacts(void* h, int i, int j, unsigned* num_found);
int orc_solve_contacts(void* h);
int orc_get_contacts(void* h, double* out, unsigned capacity, unsigned* n);

/* Named per-particle arrays in working precision (vectors are packed xyz):
 * x v a density dfsph_factor kappa kappa_v density_adv pressure
 * last_pressure aii dii dij_pj adv_density pressure_accel density_err
 * num_neighbors(u32) boundary(4/elt, [B,N]) boundary_kernel(4/elt, [B,N])
 * force torque ([B,N] vec). */
int orc_set_array(void* h, const char* name, const void* data);
int orc_get_array(void* h, const char* name, void* out);

/* Solver::update_particle_neighbors<wrap> (solver.hpp:225-260): grid,
 * neighbour list and density.  Requires boundaries computed first. */
int orc_compute_all_boundaries(void* h);
int orc_update_particle_neighbors(void* h);
/* Neighbour list in the canonical order (27 cells in the reference order,
 * ascending particle id inside a cell): ids [N, Kmax], displacements
 * [N, Kmax, 3] in working precision. */
int orc_get_neighbors(void* h, unsigned* ids, void* xij);
/* density-change sum of every particle into density_adv (needs the
 * boundary records and neighbour list of the current state) */
int orc_compute_density_change(void* h);
/* x += dt v, wrapped on periodic axes */
int orc_integrate_velocity(void* h);
/* Planted mistakes for the oracle's own tests: a non-zero code switches one
 * deliberate error into the algorithm (see enum Plant in alv_oracle.cpp),
 * process-wide, until orc_plant(0). */
int orc_plant(int code);
/* compute_dfsph_factor (runner.hpp:3124-3154) on the current lists and
 * boundary records, into dfsph_factor */
int orc_compute_dfsph_factor(void* h);
/* The velocity change of one stiffness pass (the shared body of
 * warm_start_divergence_solve_1 / divergence_solve_iteration
 * (divergence_form != 0, runner.hpp:3252-3363) or warm_start_pressure_solve1
 * / pressure_solve_iteration (runner.hpp:3452-3588)) for the stiffness field
 * k[n] (FP64 values), applied to v with the current dt; boundary forces
 * accumulate as in the step. */
int orc_apply_stiffness(void* h, int divergence_form, const double* k);
/* compute_pressure_accels (runner.hpp:2774-2817) from the current pressure,
 * into pressure_accel, without integrating */
int orc_compute_pressure_accels(void* h);
int orc_get_grid(void* h, unsigned* cell_count, unsigned* cell_start,
                 unsigned* anomaly);
int orc_step(void* h);
int orc_get_report(void* h, orc_report* r);

/* Particle-sampled boundaries (Pile with VolumeMethod::pellets,
 * pile.hpp:16,217-290; SolverI with pellets, solver_i.hpp:79-330).  After
 * orc_enable_pellets (which reallocates every array) the collision vertices
 * of each body (orc_set_collision) are its pellets, in the body frame;
 * pellet k of their concatenation is particle max_num_particles + k of x, v,
 * pressure, last_pressure, num_neighbors, num_pellet_neighbors and of the
 * lists.  orc_refresh_pellets checks the pellet count against the limit
 * (-4 when over).  orc_transform_pellets and orc_find_contacts_pellets use
 * the current body kinematics as the step start's.  orc_get_lists:
 * pellet_list 0 = fluid (bead) lists, 1 = pellet lists, [max_n + max_p][K]. */
int orc_enable_pellets(void* h, unsigned max_num_pellets);
int orc_refresh_pellets(void* h, unsigned* num_pellets);
int orc_hint_identical_sequence(void* h, unsigned begin, unsigned end);
int orc_transform_pellets(void* h);
int orc_find_contacts_pellets(void* h, unsigned* num_found);
/* MLS extrapolation onto the pellets from pressure (last = 0) or
 * last_pressure (last = 1) (runner.hpp:2543-2607) */
int orc_extrapolate_pressure(void* h, int last);
int orc_get_lists(void* h, int pellet_list, unsigned* ids, void* xij);
/* make_bead_pellet_neighbor_list on samples offset .. offset+ns-1 of
 * sample_x [offset+ns][3] against the last grid; outputs [offset+ns][K]
 * (xixj packed xyz), rows indexed by the sample's index. */
int orc_sample_bead_pellet(void* h, const double* sample_x, unsigned ns, unsigned offset,
                           unsigned max_beads, unsigned* bead_ids, double* bead_xixj,
                           unsigned* bead_counts, unsigned* pel_ids, double* pel_xixj,
                           unsigned* pel_counts);
/* List sums of n rows of K entries: what 0 = boundary kernel with pellets
 * (out [n][4]), 1 = density with pellets, 2 = DFSPH factor with pellets
 * (out [n]).  bead_xixj may be null for what 0.  The DFSPH factor with
 * pellets (runner.hpp:3156-3190) takes the pellet entries' vectors as
 * grad_wvol, as the reference does (pinned by equivalence: pellet entries
 * V gradW(r) give the factor of fluid neighbours at r). */
/* Sampling with pellet lists (what 0 density, 1 velocity, 2 vorticity;
 * runner.hpp:2122-2147, 4242-4359): lists [ns][K] (xixj packed xyz),
 * density and v indexed by particle id (v [max_n + max_p][3]), sample_v
 * [ns][3] for the vorticity; out [ns] or [ns][3]. */
int orc_sample_with_pellets(void* h, int what, const unsigned* ids, const double* xixj,
                            const unsigned* counts, const unsigned* pids, const double* pxixj,
                            const unsigned* pcounts, unsigned ns, const double* density,
                            const double* v, const double* sample_v, double* out);
int orc_pellet_sums(void* h, int what, unsigned K, unsigned n, const double* bead_xixj,
                    const unsigned* bead_counts, const double* pel_xixj,
                    const unsigned* pel_counts, double* out);

/* Stand-alone helpers used by the tests. */
int orc_peirce_constants(double* cosphi, double* sinphi, double* b,
                         double* r, double* c);
/* Cubic spline W and grad W (runner.hpp:596-638) in precision `precision`:
 * out[i*4 + 0..2] = gradW(d_i), out[i*4 + 3] = W(|d_i|) via d2 form. */
int orc_cubic_kernel(int precision, double h, size_t n, const double* d,
                     double* out);
/* 32-node cubic Lagrange shape functions and gradients
 * (runner.hpp:1259-1584) at local coordinate xi (double). */
int orc_shape_functions(const double* xi, double* N, double* dN);
/* Node position of node index l (runner.hpp:1131-1178). */
int orc_node_position(const double* domain_min, const unsigned* res,
                      const double* cell_size, unsigned l, double* out);
/* Positions of all nodes (out[3 * num_nodes]). */
int orc_node_positions(const double* domain_min, const unsigned* res,
                       const double* cell_size, double* out);
/* Indices of the 32 nodes of cell ipos (runner.hpp:1197-1257). */
int orc_cell_nodes(const unsigned* res, const int* ipos, unsigned* out);
/* Volume map from a sampled distance map (the grid variant,
 * runner.hpp:1616-1658), precision 4 or 8, nodes [begin, end). */
int orc_volume_from_distance_map(int precision, double kernel_radius,
                                 const double* domain_min,
                                 const double* domain_max,
                                 const unsigned* res, const double* cell_size,
                                 double sign, const double* distance_nodes,
                                 unsigned begin, unsigned end, double* out);
/* Serial FP64 sum (Runner::sum with FP64 accumulation). */
double orc_sum_f32(const float* v, size_t n);

/* Learned warm-start network (mlp_oracle.cpp): 8 -> 64 -> 64 -> 64 -> 2,
 * SiLU, flat FP32 parameters W1 b1 W2 b2 W3 b3 W4 b4 (9026).  `quantize`
 * selects the bfloat16 rounding points of the device kernels; 0 is exact
 * FP64.  x: [n][8] normalised inputs, t: [n][2] normalised targets. */
int orc_mlp_num_params(void);
int orc_mlp_forward(const float* params, size_t n, const float* x, int quantize, double* y);
int orc_mlp_loss_grad(const float* params, size_t n, const float* x, const float* t,
                      int quantize, double* loss, double* grad);

#ifdef __cplusplus
}
#endif
#endif
