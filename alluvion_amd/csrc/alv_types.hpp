// alv_types.hpp -- plain-old-data blocks shared by the host runtime and the
// HIP kernels.  SimParams replaces the reference's process-global
// __constant__ Const<TF>/Consti (constants.hpp:40-143): it is built per
// solver on the host and passed to every kernel by value, so several solvers
// (or scenes) with different constants can coexist in one process.
#pragma once

#include <stdint.h>

namespace alv {

// ALV_DIST_MAP: a signed distance sampled on the boundary's map nodes (the
// reference's grid path, used for MeshDistance)
enum { ALV_DIST_BOX = 0, ALV_DIST_BOX_SHELL = 1, ALV_DIST_SPHERE = 2, ALV_DIST_MAP = 3 };

struct SimParams {
  // kernel constants (set_kernel_radius, constants.hpp:100-120)
  float h, h2, k, l, w0;
  // correctly rounded 1/h and 1/h^2 (divisions by h and h^2 use div_known)
  float inv_h, inv_h2;
  // particle attributes (set_particle_attr, constants.hpp:122-128)
  float radius, vol, mass, rho0;
  float visc, bvisc, eps;
  float g[3];
  // periodic wrap (set_wrap_length, constants.hpp:130-136); bit a = axis a
  int wrap_mask;
  float wrap_len[3], wrap_min[3], wrap_max[3];
  // neighbour grid (Consti, constants.hpp:40-49)
  int res[3], off[3];
  unsigned kmax, cell_cap, num_cells;
  unsigned num_boundaries;
  unsigned n;  // particles in this step
  // Where this scene sits in the solver's arrays (one solver can hold many
  // independent scenes): particles [first, first + n) of a range of `span`
  // slots, a multiple of the 256-thread block, so that no block straddles two
  // scenes; cells [cell_base, cell_base + num_cells] (the last one is the bin
  // of particles outside the grid); boundaries from bnd_base.
  unsigned first, span, cell_base, bnd_base;
  // particle-sampled boundaries (VolumeMethod::pellets): the boundaries are
  // the np pellets instead of volume maps
  int pellet_mode;
  unsigned np;
};

// Geometry of one 32-node cubic Lagrange map (Pile lists, pile.hpp:106-112).
struct GridGeomDev {
  float dmin[3], dmax[3], cs[3];
  unsigned res[3];
  unsigned num_nodes;
};

// One boundary (container or rigid body) as the kernels see it.
struct BoundaryDev {
  int type;
  float sign;
  float half[3], half_outer[3];
  float outset, radius, max_distance;
  // square of a radius around the body origin that holds the map's box (a
  // point farther from x lies outside the map)
  float map_radius2;
  GridGeomDev geom;
  const float* volume;  // device pointer to the volume map nodes
  const float* distance;  // ALV_DIST_MAP: device pointer to the distance map nodes
  // per map cell (x fastest): 1 when any of its 32 nodes is non-zero; a
  // point in a cell with 0 has volume exactly 0
  const unsigned char* cell_nz;
  float x[3], v[3], omega[3];
  float q[4];  // (x, y, z, w)
  float mass;
  int dynamic;  // mass != 0: the body moves and its force sums are reduced
  // per-particle forces and torques are accumulated for it: dynamic bodies,
  // and every boundary when the solver records static boundaries' forces
  int accum;
};

// Device-resident solver control block: time step, convergence bookkeeping
// and report scalars.  Kernels read dt from here, so a step needs no host
// round trip to pick its time step.
struct StepCtl {
  float dt;
  float cfl, min_dt, max_dt;
  float tol_div;   // density_change_tolerance
  float tol_dens;  // density_error_tolerance
  unsigned min_div, max_div, min_dens, max_dens;
  unsigned div_iter, dens_iter;
  unsigned div_done, dens_done;
  float mean_density_change, mean_density_error;
  float particle_max_v2, pile_max_v2, max_v2, cfl_dt, utilized_cfl;
  float length_scale;
  float div_dt;  // the time step of the step's divergence solve (set by update_dt)
  unsigned pad;
};

}  // namespace alv
