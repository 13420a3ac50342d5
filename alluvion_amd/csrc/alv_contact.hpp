// alv_contact.hpp -- rigid contact detection: job table and launcher.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/alv/alv.h"
#include "alv_types.hpp"

namespace alv {

// One (i, j) body pair: the collision vertices of body i (vbase .. vbase+nv
// in the pile's vertex array) against body j's distance.  Candidates are
// numbered pair by pair, cbase being the first candidate of this pair.
struct ContactJob {
  unsigned i, j;
  unsigned vbase, nv, cbase;
  float restitution;  // restitution_i * restitution_j
  float friction;     // friction_i + friction_j
  unsigned pad;
};

void launch_collision(const ContactJob* jobs, unsigned num_jobs, unsigned num_candidates,
                      const BoundaryDev* bodies, const float* inertia, const float* verts,
                      float tolerance, unsigned max_contacts, alv_contact* contacts,
                      unsigned* count, hipStream_t s);

}  // namespace alv
