// alv_contact.hpp -- rigid contact detection: job table and launcher.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/alv/alv.h"
#include "alv_types.hpp"

namespace alv {

// One body i against a run of nj consecutive bodies j .. j+nj-1 (nj = 1 for
// a plain (i, j) pair): the collision vertices vbase .. vbase+nv of the
// pile's vertex array.  Candidates are numbered job by job, cbase being the
// first of this job, and inside a job vertex-major: (vertex, j).
enum : unsigned {
  kJobWorld = 1u,  // vertices are world positions (pellets) rather than body-frame
  kJobSkip = 2u,   // reject i == j and pairs of two massless bodies per candidate
};
struct ContactJob {
  unsigned i, j;
  unsigned vbase, nv, cbase;
  unsigned nj;
  unsigned flags;
  unsigned pad;
};

// inertia: [B][3] principal inertia, then [B][2] (restitution, friction)
void launch_collision(const ContactJob* jobs, unsigned num_jobs, unsigned num_candidates,
                      unsigned nb, const BoundaryDev* bodies, const float* inertia,
                      const float* verts, float tolerance, unsigned max_contacts,
                      alv_contact* contacts, unsigned* count, hipStream_t s);

}  // namespace alv
