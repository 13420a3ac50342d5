// alv_capi.hip -- implementation of the C ABI (include/alv/alv.h): device
// memory, boundary (Pile) management, volume-map construction and the
// per-step orchestration of the DFSPH / IISPH solvers.
//
// Orchestration differs from the reference's Runner (runner.hpp:4817-4902),
// which synchronises after every launch: a step here is a stream of
// asynchronous launches on one HIP stream.  The Jacobi loops are enqueued in
// chunks; iteration kernels check a device flag and return at once after
// convergence, so the host waits once per chunk rather than once per
// iteration, and once at the end of the step to read the report.

#include <float.h>
#include <stdio.h>
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <new>
#include <string>
#include <memory>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/alv/alv.h"
#include "alv_device.hpp"
#include "alv_contact.hpp"
#include "alv_kernels.hpp"

using namespace alv;

namespace {
thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}
}  // namespace

namespace alv {
// error reporting shared with the other C ABI translation units
int set_error(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace alv

namespace {

#define ALV_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail(ALV_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
int dev_alloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  ALV_HIP(hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T)));
  ALV_HIP(hipMemset(*p, 0, count * sizeof(T)));
  return ALV_OK;
}

// ---------------------------------------------------------------------------
// Peirce product rule (M = 14): cos(phi), sin(phi), b from the 30-point
// Gauss-Legendre rule; r = sqrt((x+1)/2), c = w/(4 sqrt 2) from the 15-point
// Gauss-Jacobi(0, 1/2) rule.  Nodes and weights come from the Golub-Welsch
// eigenproblem of the Jacobi matrix (implicit-shift QL, long double).
// ---------------------------------------------------------------------------
void gauss_rule(int n, const std::vector<long double>& a, const std::vector<long double>& b,
                long double mu0, std::vector<long double>& x, std::vector<long double>& w) {
  std::vector<long double> d(a), e(n, 0.0L), z(n, 0.0L);
  for (int i = 0; i + 1 < n; ++i) e[i] = b[i];
  z[0] = 1.0L;
  for (int l = 0; l < n; ++l) {
    for (int guard = 0; guard < 100; ++guard) {
      int m = l;
      for (; m < n - 1; ++m)
        if (fabsl(e[m]) <= 1e-30L * (fabsl(d[m]) + fabsl(d[m + 1])) + 1e-300L) break;
      if (m == l) break;
      long double g = (d[l + 1] - d[l]) / (2.0L * e[l]);
      long double r = hypotl(g, 1.0L);
      g = d[m] - d[l] + e[l] / (g + copysignl(r, g));
      long double s = 1.0L, c = 1.0L, p = 0.0L;
      bool deflated = false;
      for (int i = m - 1; i >= l; --i) {
        long double f = s * e[i], bb = c * e[i];
        r = hypotl(f, g);
        e[i + 1] = r;
        if (r == 0.0L) {
          d[i + 1] -= p;
          e[m] = 0.0L;
          deflated = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0L * c * bb;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - bb;
        long double t = z[i + 1];
        z[i + 1] = s * z[i] + c * t;
        z[i] = c * z[i] - s * t;
      }
      if (deflated) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0L;
    }
  }
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::sort(idx.begin(), idx.end(), [&](int i, int j) { return d[i] < d[j]; });
  x.resize(n);
  w.resize(n);
  for (int i = 0; i < n; ++i) {
    x[i] = d[idx[i]];
    w[i] = mu0 * z[idx[i]] * z[idx[i]];
  }
}

void peirce_table(float out[120]) {
  std::vector<long double> x, w;
  {
    const int n = 30;
    std::vector<long double> a(n, 0.0L), b(n - 1);
    for (int k = 1; k < n; ++k) b[k - 1] = sqrtl((long double)k * k / (4.0L * k * k - 1.0L));
    gauss_rule(n, a, b, 2.0L, x, w);
    for (int j = 0; j < n; ++j) {
      out[j] = float(x[j]);
      out[30 + j] = float(sqrtl(1.0L - x[j] * x[j]));
      out[60 + j] = float(w[j]);
    }
  }
  {
    const int n = 15;
    const long double be = 0.5L;
    std::vector<long double> a(n), b(n - 1);
    for (int k = 0; k < n; ++k) {
      long double s = 2.0L * k + be;
      a[k] = (be * be) / (s * (s + 2.0L));
    }
    for (int k = 1; k < n; ++k) {
      long double s = 2.0L * k + be;
      b[k - 1] = sqrtl(4.0L * k * k * (k + be) * (k + be) / (s * s * (s + 1.0L) * (s - 1.0L)));
    }
    gauss_rule(n, a, b, powl(2.0L, 2.5L) / 3.0L, x, w);
    for (int k = 0; k < n; ++k) {
      out[90 + k] = float(sqrtl((x[k] + 1.0L) / 2.0L));
      out[105 + k] = float(w[k] / (4.0L * sqrtl(2.0L)));
    }
  }
}

SimParams make_params(const alv_const& cn, const alv_consti& cni, unsigned num_boundaries,
                      unsigned n) {
  SimParams p;
  memset(&p, 0, sizeof(p));
  p.h = cn.kernel_radius;
  p.h2 = cn.kernel_radius_sqr;
  // IEEE single-precision divisions on the host: RN(1/h), RN(1/h^2)
  p.inv_h = 1.0f / p.h;
  p.inv_h2 = 1.0f / p.h2;
  p.k = cn.cubic_kernel_k;
  p.l = cn.cubic_kernel_l;
  p.w0 = cn.cubic_kernel_zero;
  p.radius = cn.particle_radius;
  p.vol = cn.particle_vol;
  p.mass = cn.particle_mass;
  p.rho0 = cn.density0;
  p.visc = cn.viscosity;
  p.bvisc = cn.boundary_viscosity;
  p.eps = cn.dfsph_factor_epsilon;
  for (int a = 0; a < 3; ++a) {
    p.g[a] = cn.gravity[a];
    float L = cn.wrap_length[a];
    p.wrap_len[a] = L;
    p.wrap_min[a] = L * -0.5f;
    p.wrap_max[a] = L * 0.5f;
    if (L > 0.0f) p.wrap_mask |= 1 << a;
    p.res[a] = int(cni.grid_res[a]);
    p.off[a] = cni.grid_offset[a];
  }
  p.kmax = cni.max_num_neighbors_per_particle;
  p.cell_cap = cni.max_num_particles_per_cell;
  p.num_cells = cni.grid_res[0] * cni.grid_res[1] * cni.grid_res[2];
  p.num_boundaries = num_boundaries;
  p.n = n;
  return p;
}

// Per-kernel timing: HIP events recorded around every launch on the launch
// stream; elapsed times are collected after the step's final synchronisation.
class EventProfiler : public KernelProfiler {
 public:
  struct Stat {
    std::string name;
    double total_ms = 0.0;
    unsigned long long launches = 0;
  };
  ~EventProfiler() override {
    for (auto& p : pending_) {
      (void)hipEventDestroy(p.b);
      (void)hipEventDestroy(p.e);
    }
    for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
  }
  void begin(const char* name, hipStream_t s) override {
    Pending p;
    p.name = name;
    p.b = get();
    p.e = get();
    (void)hipEventRecord(p.b, s);
    pending_.push_back(p);
  }
  void end(const char*, hipStream_t s) override {
    if (!pending_.empty()) (void)hipEventRecord(pending_.back().e, s);
  }
  void collect() {
    for (auto& p : pending_) {
      float ms = 0.f;
      if (hipEventSynchronize(p.e) == hipSuccess && hipEventElapsedTime(&ms, p.b, p.e) == hipSuccess) {
        Stat& st = find(p.name);
        st.total_ms += ms;
        st.launches += 1;
      }
      pool_.push_back(p.b);
      pool_.push_back(p.e);
    }
    pending_.clear();
  }
  void reset() {
    collect();
    stats_.clear();
  }
  std::vector<Stat>& stats() { return stats_; }

 private:
  struct Pending {
    const char* name;
    hipEvent_t b, e;
  };
  hipEvent_t get() {
    if (!pool_.empty()) {
      hipEvent_t e = pool_.back();
      pool_.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
  }
  Stat& find(const char* name) {
    for (auto& s : stats_)
      if (s.name == name) return s;
    stats_.push_back(Stat());
    stats_.back().name = name;
    return stats_.back();
  }
  std::vector<Pending> pending_;
  std::vector<hipEvent_t> pool_;
  std::vector<Stat> stats_;
};
EventProfiler* g_event_profiler = nullptr;
}  // namespace

// ---------------------------------------------------------------------------
// handles
// ---------------------------------------------------------------------------
struct alv_pile {
  std::vector<BoundaryDev> bnd;  // kinematics live in bnd[b].x/v/q/omega
  std::vector<float*> volumes;
  std::vector<float*> distances;  // distance maps (ALV_DISTANCE_MAP), else null
  std::vector<unsigned char*> cell_nz;  // non-zero flags of the map cells
  // set when a caller got a writable pointer to a volume map: the flags are
  // recomputed before the next step
  mutable bool maps_dirty = false;
  std::vector<float> restitution, friction;
  std::vector<std::array<float, 3>> inertia, force, torque, accel;
  float gravity[3] = {0.f, 0.f, 0.f};
  unsigned max_num_contacts = 0;
  float* peirce = nullptr;  // device, 120 floats
  // rigid contacts: collision vertices per body (host, body frame), their
  // concatenation on the device, and the contact buffers
  std::vector<std::vector<float>> coll_verts;
  bool verts_dirty = true;
  float* verts_dev = nullptr;
  size_t verts_cap = 0;  // floats
  unsigned* vbody_dev = nullptr;  // body of every collision vertex (pellet)
  size_t vbody_cap = 0;
  ContactJob* jobs_dev = nullptr;
  size_t jobs_cap = 0;
  BoundaryDev* bodies_dev = nullptr;
  size_t bodies_cap = 0;
  float* inertia_dev = nullptr;
  size_t inertia_cap = 0;
  alv_contact* contacts_dev = nullptr;
  unsigned* count_dev = nullptr;
  std::vector<alv_contact> contacts;  // host copy of the kept contacts
  unsigned num_found = 0;
  // device kinematics (x, v, q, omega of every body) as of the last copy:
  // the snapshot is taken on the host at each step start and uploaded when
  // the device copy is asked for
  std::vector<float> kin_snap;
  bool kin_dirty = false;
  // VolumeMethod::pellets (pile.hpp:16,124-130): the collision vertices of
  // every body are its pellets; no volume maps are built
  int volume_method = ALV_VOLUME_MAP;
  unsigned max_pellets = 0;
  std::vector<std::pair<unsigned, unsigned>> ident;  // hint_identical_sequence runs
  float* kin_dev = nullptr;
  size_t kin_cap = 0;
  ~alv_pile() {
    for (float* v : volumes) (void)hipFree(v);
    for (float* d : distances) (void)hipFree(d);
    for (unsigned char* f : cell_nz) (void)hipFree(f);
    if (peirce) (void)hipFree(peirce);
    if (verts_dev) (void)hipFree(verts_dev);
    if (vbody_dev) (void)hipFree(vbody_dev);
    if (jobs_dev) (void)hipFree(jobs_dev);
    if (bodies_dev) (void)hipFree(bodies_dev);
    if (inertia_dev) (void)hipFree(inertia_dev);
    if (contacts_dev) (void)hipFree(contacts_dev);
    if (count_dev) (void)hipFree(count_dev);
    if (kin_dev) (void)hipFree(kin_dev);
  }
  // x [B][3], v [B][3], q [B][4], omega [B][3]
  void snapshot_kinematics() {
    const size_t B = bnd.size();
    kin_snap.assign(B * 13, 0.f);
    for (size_t b = 0; b < B; ++b)
      for (int a = 0; a < 4; ++a) {
        if (a < 3) {
          kin_snap[b * 3 + a] = bnd[b].x[a];
          kin_snap[B * 3 + b * 3 + a] = bnd[b].v[a];
          kin_snap[B * 10 + b * 3 + a] = bnd[b].omega[a];
        }
        kin_snap[B * 6 + b * 4 + a] = bnd[b].q[a];
      }
    kin_dirty = true;
  }
};

struct alv_solver {
  int kind = ALV_SOLVER_DF;
  alv_pile* pile = nullptr;
  unsigned cap = 0;
  // particle-sampled boundaries: pellet rows after the fluid's, and the
  // launch table of the pellets (one scene, the fluid's constants)
  unsigned pcap = 0;
  SimParams* ppar_dev = nullptr;
  unsigned short* pblk_scene = nullptr;
  alv_const cn;
  alv_consti cni;
  alv_solver_settings set;
  alv_solver_report rep;
  DevArrays A;
  hipStream_t stream = nullptr;
  // Step inputs: the control blocks [NS] followed by the boundaries
  // [bnd_capacity], in one pinned, coherent host block and one device block,
  // so one copy uploads both; the kernels also publish the control blocks
  // back into the host block (publish_ctl).
  void* step_host = nullptr;
  void* step_dev = nullptr;
  StepCtl* ctl_host = nullptr;  // views into step_host / step_dev
  BoundaryDev* bnd_host = nullptr;
  BoundaryDev* bnd_dev = nullptr;
  double* body_host = nullptr;  // pinned [B*6]
  unsigned bnd_capacity = 0;
  unsigned num_cells = 0;
  unsigned dirty = 0;
  unsigned last_div = 2, last_dens = 3;
  unsigned* export_ids = nullptr;
  unsigned* export_counts = nullptr;
  // Reference-layout copies made on request (user order, owned here):
  // particle_neighbors [cap][K] float4, particle_boundary /
  // particle_boundary_kernel [B][cap] float4, particle_force /
  // particle_torque [B][cap] float3, and the IISPH arrays.
  float4* export_nbr = nullptr;
  float4* export_brec = nullptr;
  float4* export_bker = nullptr;
  float* export_force = nullptr;
  float* export_torque = nullptr;
  float* export_ii_scalar[4] = {};  // last_pressure, aii, adv_density, density_err
  float* export_ii_vec[3] = {};  // dii, dij_pj, pressure_accel
  unsigned export_stale = ~0u;  // kX* bits of the copies older than the state
  int record_static_forces = 0;  // accumulate particle forces of static boundaries too
  alv_mlp* net = nullptr;  // learned warm start (not owned)
  int net_div = 0, net_dens = 0;  // which warm starts use it
  int log_features = 0;
  int features_valid = 0;  // the last step recorded warm-start features
  // Scenes held by this solver: one for a solver made by alv_solver_create,
  // several for a batch.  Every scene owns a block-aligned range of particle
  // slots, a cell range and the boundaries of its pile.
  struct Scene {
    alv_pile* pile;
    unsigned n, first, span;
    unsigned cell_base, num_cells;
    int off[3];
    unsigned res[3];
    unsigned bnd_base;
    float t, dt;  // batch: time and step of the scene
    unsigned last_div, last_dens;
    alv_solver_report rep;
  };
  std::vector<Scene> scenes;
  bool batch = false;
  unsigned cells_total = 0;  // cell slots of all scenes
  SimParams* par_dev = nullptr;
  SimParams* par_host = nullptr;  // pinned; the table as last uploaded
  int par_valid = 0;
  unsigned short* blk_scene = nullptr;
  unsigned short* cblk_scene = nullptr;
  // Step segments replayed as HIP graphs (captured on first use and again
  // when what they contain changes).
  struct Segment {
    hipGraphExec_t exec = nullptr;
    unsigned long long key = ~0ull;
  };
  Segment seg[3];
  int use_graph = 0;
  int grid_valid = 0;  // a neighbour search has filled the grid (sampling needs it)
  unsigned slack = 0;  // Jacobi iterations enqueued beyond the last step's count
  unsigned chunk_div = 0, chunk_dens = 0, shrink_div = 0, shrink_dens = 0;
  std::vector<void*> allocations;

  ~alv_solver() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (void* p : allocations) (void)hipFree(p);
    if (export_ids) (void)hipFree(export_ids);
    if (export_counts) (void)hipFree(export_counts);
    for (void* p : {(void*)export_nbr, (void*)export_brec, (void*)export_bker,
                    (void*)export_force, (void*)export_torque})
      if (p) (void)hipFree(p);
    for (float* p : export_ii_scalar)
      if (p) (void)hipFree(p);
    for (float* p : export_ii_vec)
      if (p) (void)hipFree(p);
    if (step_dev) (void)hipFree(step_dev);
    if (step_host) (void)hipHostFree(step_host);
    if (body_host) (void)hipHostFree(body_host);
    if (par_host) (void)hipHostFree(par_host);
    for (auto& g : seg)
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (stream) (void)hipStreamDestroy(stream);
  }
  template <class T>
  int alloc(T** p, size_t count) {
    int rc = dev_alloc(p, count);
    if (rc == ALV_OK) allocations.push_back(*p);
    return rc;
  }
};

// ---------------------------------------------------------------------------
// device and memory
// ---------------------------------------------------------------------------
extern "C" {

const char* alv_last_error(void) { return g_last_error.c_str(); }
const char* alv_version(void) { return "alluvion-amd 0.1 (gfx950)"; }

int alv_set_device(int device) {
  ALV_HIP(hipSetDevice(device));
  return ALV_OK;
}
int alv_get_device_count(int* count) {
  if (!count) return fail(ALV_ERR_ARG, "count is null");
  ALV_HIP(hipGetDeviceCount(count));
  return ALV_OK;
}
int alv_device_synchronize(void) {
  ALV_HIP(hipDeviceSynchronize());
  return ALV_OK;
}
int alv_malloc(void** dptr, size_t bytes) {
  if (!dptr) return fail(ALV_ERR_ARG, "dptr is null");
  *dptr = nullptr;
  if (bytes == 0) return ALV_OK;
  ALV_HIP(hipMalloc(dptr, bytes));
  ALV_HIP(hipMemset(*dptr, 0, bytes));
  return ALV_OK;
}
int alv_free(void* dptr) {
  if (dptr) ALV_HIP(hipFree(dptr));
  return ALV_OK;
}
int alv_malloc_pinned(void** hptr, size_t bytes) {
  if (!hptr) return fail(ALV_ERR_ARG, "hptr is null");
  *hptr = nullptr;
  if (bytes == 0) return ALV_OK;
  ALV_HIP(hipHostMalloc(hptr, bytes, hipHostMallocDefault));
  memset(*hptr, 0, bytes);
  return ALV_OK;
}
int alv_free_pinned(void* hptr) {
  if (hptr) ALV_HIP(hipHostFree(hptr));
  return ALV_OK;
}
int alv_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  if (bytes) ALV_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return ALV_OK;
}
int alv_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  if (bytes) ALV_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return ALV_OK;
}
int alv_memcpy_d2d(void* dst, const void* src, size_t bytes) {
  if (bytes) ALV_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  return ALV_OK;
}
int alv_memset(void* dptr, int value, size_t bytes) {
  if (bytes) ALV_HIP(hipMemset(dptr, value, bytes));
  return ALV_OK;
}

}  // extern "C"

// reductions for Runner::sum / Runner::max
__global__ __launch_bounds__(256) void k_sum_f32(const float* d, size_t n, double* part) {
  __shared__ double lds[4];
  double acc = 0.0;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x)
    acc += double(d[i]);
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  if ((threadIdx.x & 63u) == 0) lds[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = lds[0] + lds[1] + lds[2] + lds[3];
}
__global__ __launch_bounds__(256) void k_max_f32(const float* d, size_t n, float* part) {
  __shared__ float lds[4];
  float acc = -FLT_MAX;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x)
    acc = fmaxf(acc, d[i]);
  for (int o = 32; o > 0; o >>= 1) acc = fmaxf(acc, __shfl_down(acc, o, 64));
  if ((threadIdx.x & 63u) == 0) lds[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = fmaxf(fmaxf(lds[0], lds[1]), fmaxf(lds[2], lds[3]));
}

// Final step of the allocation-free sum: one lane adds the block partials in
// index order, the order in which alv_sum_f32 adds them on the host, so both
// entry points give the same bits.
__global__ __launch_bounds__(64) void k_sum_partials(double* part, unsigned nb) {
  if (threadIdx.x != 0) return;
  double s = 0.0;
  for (unsigned k = 0; k < nb; ++k) s += part[k];
  part[nb] = s;
}

// Scratch of alv_sum_f32_noalloc: the block partials and the result.  Every
// GPU the library runs on holds its own copy until the process ends.  Host
// code must never pass `scratch` itself to a kernel or a copy (on the host it
// names a stand-in, not device memory); it uses the per-device address from
// hipGetSymbolAddress (scratch_address below).
static __device__ __align__(256) double scratch[ALV_SUM_NOALLOC_SCRATCH_DOUBLES];

namespace {
constexpr unsigned kSumBlocks = 1024;  // block partials of k_sum_f32
static_assert(ALV_SUM_NOALLOC_SCRATCH_DOUBLES >= kSumBlocks + 1, "scratch too small");

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device with that device current.
int scratch_address(double** ptr) {
  int dev = 0;
  ALV_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(ALV_ERR_UNSUPPORTED, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    ALV_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(ALV_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *ptr = g_scratch_ptr[dev];
  return ALV_OK;
}
}  // namespace

extern "C" {

int alv_sum_f32(const float* d, size_t n, double* out) {
  if (!out) return fail(ALV_ERR_ARG, "out is null");
  const unsigned nb = kSumBlocks;
  double* part = nullptr;
  ALV_HIP(hipMalloc(&part, nb * sizeof(double)));
  hipLaunchKernelGGL(k_sum_f32, dim3(nb), dim3(256), 0, 0, d, n, part);
  std::vector<double> h(nb);
  hipError_t e = hipMemcpy(h.data(), part, nb * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(part);
  if (e != hipSuccess) return fail(ALV_ERR_HIP, hipGetErrorString(e));
  double s = 0.0;
  for (double v : h) s += v;
  *out = s;
  return ALV_OK;
}

int alv_sum_f32_noalloc(const float* d, size_t n, double* out) {
  if (!out) return fail(ALV_ERR_ARG, "out is null");
  if (n && !d) return fail(ALV_ERR_ARG, "d is null");
  if (n > ALV_SUM_NOALLOC_MAX_ELEMENTS)
    return fail(ALV_ERR_ARG, "alv_sum_f32_noalloc: n = " + std::to_string(n) +
                                 " exceeds the maximum of " +
                                 std::to_string((unsigned long long)ALV_SUM_NOALLOC_MAX_ELEMENTS) +
                                 " elements");
  double* part = nullptr;
  int rc = scratch_address(&part);
  if (rc != ALV_OK) return rc;
  // Every block writes its partial on every call, so nothing left by an
  // earlier call is read.
  hipLaunchKernelGGL(k_sum_f32, dim3(kSumBlocks), dim3(256), 0, 0, d, n, part);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, 0, part, kSumBlocks);
  ALV_HIP(hipGetLastError());
  ALV_HIP(hipMemcpy(out, part + kSumBlocks, sizeof(double), hipMemcpyDeviceToHost));
  return ALV_OK;
}
int alv_max_f32(const float* d, size_t n, float* out) {
  if (!out) return fail(ALV_ERR_ARG, "out is null");
  const unsigned nb = 1024;
  float* part = nullptr;
  ALV_HIP(hipMalloc(&part, nb * sizeof(float)));
  hipLaunchKernelGGL(k_max_f32, dim3(nb), dim3(256), 0, 0, d, n, part);
  std::vector<float> h(nb);
  hipError_t e = hipMemcpy(h.data(), part, nb * sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(part);
  if (e != hipSuccess) return fail(ALV_ERR_HIP, hipGetErrorString(e));
  float m = -FLT_MAX;
  for (float v : h) m = std::max(m, v);
  *out = m;
  return ALV_OK;
}

// ---------------------------------------------------------------------------
// constants
// ---------------------------------------------------------------------------
void alv_const_set_kernel_radius(alv_const* cn, float h) {
  const float pi = 3.14159265358979323846f;
  float h2 = h * h;
  float h3 = h2 * h;
  float h4 = h2 * h2;
  cn->kernel_radius = h;
  cn->kernel_radius_sqr = h2;
  cn->cubic_kernel_k = 8.0f / h3 / pi;
  cn->cubic_kernel_l = 48.0f / h4 / pi;
  cn->cubic_kernel_zero = cn->cubic_kernel_k;
}
void alv_const_set_particle_attr(alv_const* cn, float radius, float mass, float density) {
  float vol = mass / density;
  cn->particle_radius = radius;
  cn->particle_mass = mass;
  cn->density0 = density;
  cn->particle_vol = vol;
}

// ---------------------------------------------------------------------------
// Pile
// ---------------------------------------------------------------------------
int alv_pile_create(alv_pile** out, unsigned max_num_contacts, int volume_method) {
  return alv_pile_create_ex(out, max_num_contacts, volume_method, 10000u);
}

int alv_pile_create_ex(alv_pile** out, unsigned max_num_contacts, int volume_method,
                       unsigned max_num_pellets) {
  if (!out) return fail(ALV_ERR_ARG, "out is null");
  if (volume_method != ALV_VOLUME_MAP && volume_method != ALV_PELLETS)
    return fail(ALV_ERR_ARG, "unknown volume method");
  alv_pile* p = new (std::nothrow) alv_pile();
  if (!p) return fail(ALV_ERR_STATE, "out of host memory");
  p->max_num_contacts = max_num_contacts;
  p->volume_method = volume_method;
  // max_num_pellets_ is 0 for volume maps (pile.hpp:162-163)
  p->max_pellets = volume_method == ALV_PELLETS ? max_num_pellets : 0u;
  float table[120];
  peirce_table(table);
  hipError_t e = hipMalloc(&p->peirce, sizeof(table));
  if (e == hipSuccess) e = hipMemcpy(p->peirce, table, sizeof(table), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    delete p;
    return fail(ALV_ERR_HIP, hipGetErrorString(e));
  }
  *out = p;
  return ALV_OK;
}

int alv_pile_destroy(alv_pile* pile) {
  delete pile;
  return ALV_OK;
}

}  // extern "C"

namespace {
// Pile::add and Pile::replace (pile.hpp:217-352, 353-492): builds boundary
// `slot` (appended when slot < 0) with its map domain, resolution and, for
// volume maps, its map on the device.
int place_boundary(alv_pile* pile, int slot, const alv_const* cn, const alv_distance* dist,
                   const unsigned resolution[3], float cell_width, float margin, float sign,
                   float mass, float restitution, float friction, const float inertia[3],
                   const float x[3], const float q[4], unsigned* boundary_id) {
  if (!pile || !cn || !dist) return fail(ALV_ERR_ARG, "null argument");
  if (slot < 0 && pile->bnd.size() >= 32) return fail(ALV_ERR_UNSUPPORTED, "at most 32 boundaries");
  if (slot >= 0 && size_t(slot) >= pile->bnd.size()) return fail(ALV_ERR_ARG, "bad boundary id");
  BoundaryDev B;
  memset(&B, 0, sizeof(B));
  B.type = dist->type;
  B.sign = sign;
  float aabb_min[3], aabb_max[3];
  const float* pr = dist->params;
  if (dist->type == ALV_DISTANCE_BOX) {
    // BoxDistance (dg/box_distance.hpp:11-17)
    B.outset = pr[3];
    float hx2 = 0.f;
    for (int a = 0; a < 3; ++a) {
      B.half[a] = pr[a] * 0.5f;
      aabb_min[a] = pr[a] * -0.5f - B.outset;
      aabb_max[a] = B.half[a] + B.outset;
    }
    hx2 = B.half[0] * B.half[0] + B.half[1] * B.half[1] + B.half[2] * B.half[2];
    B.max_distance = sqrtf(hx2) + B.outset;
  } else if (dist->type == ALV_DISTANCE_BOX_SHELL) {
    // BoxShellDistance (dg/box_shell_distance.hpp:11-24)
    float th = pr[3];
    B.outset = pr[4];
    for (int a = 0; a < 3; ++a) {
      B.half[a] = pr[a] * 0.5f;
      B.half_outer[a] = B.half[a] + th;
      aabb_min[a] = pr[a] * -0.5f - th - B.outset;
      aabb_max[a] = B.half[a] + th + B.outset;
    }
    float o2 = B.half_outer[0] * B.half_outer[0] + B.half_outer[1] * B.half_outer[1] +
               B.half_outer[2] * B.half_outer[2];
    B.max_distance = sqrtf(o2) + B.outset;
  } else if (dist->type == ALV_DISTANCE_SPHERE) {
    // SphereDistance (dg/sphere_distance.hpp:10-16)
    B.radius = pr[0];
    for (int a = 0; a < 3; ++a) {
      aabb_min[a] = -B.radius;
      aabb_max[a] = B.radius;
    }
    B.max_distance = B.radius;
  } else if (dist->type == ALV_DISTANCE_MAP) {
    if (!dist->signed_distance) return fail(ALV_ERR_ARG, "distance map without a function");
    if (pile->volume_method == ALV_PELLETS)
      return fail(ALV_ERR_UNSUPPORTED, "distance maps in a pellet pile");
    // the farthest corner of the bounding box from the body origin
    float r2 = 0.f;
    for (int a = 0; a < 3; ++a) {
      aabb_min[a] = dist->aabb_min[a];
      aabb_max[a] = dist->aabb_max[a];
      if (!(aabb_max[a] > aabb_min[a])) return fail(ALV_ERR_ARG, "empty bounding box");
      const float m = std::max(std::fabs(aabb_min[a]), std::fabs(aabb_max[a]));
      r2 += m * m;
    }
    B.max_distance = sqrtf(r2);
  } else {
    return fail(ALV_ERR_UNSUPPORTED, "distance type not supported");
  }
  // calculate_domain_and_resolution (pile.hpp:40-65)
  if (std::isinf(margin)) margin = sign < 0 ? cn->kernel_radius : cn->kernel_radius * 2;
  unsigned res[3] = {resolution[0], resolution[1], resolution[2]};
  for (int a = 0; a < 3; ++a) {
    B.geom.dmin[a] = aabb_min[a] - margin;
    B.geom.dmax[a] = aabb_max[a] + margin;
  }
  if (res[0] == 0 || res[1] == 0 || res[2] == 0) {
    if (cell_width == 0) cell_width = cn->particle_radius;
    for (int a = 0; a < 3; ++a)
      if (res[a] == 0) res[a] = unsigned((B.geom.dmax[a] - B.geom.dmin[a]) / cell_width);
  }
  // calculate_grid_attributes (pile.hpp:66-78)
  for (int a = 0; a < 3; ++a) {
    if (res[a] == 0) return fail(ALV_ERR_ARG, "zero grid resolution");
    B.geom.res[a] = res[a];
    B.geom.cs[a] = (B.geom.dmax[a] - B.geom.dmin[a]) / float(res[a]);
  }
  {
    // the farthest corner of the map's box from the body origin, with a
    // relative margin far above the rounding of the kernels' distance test
    double r2 = 0.0;
    for (int a = 0; a < 3; ++a) {
      const double lo = B.geom.dmin[a], hi = B.geom.dmax[a];
      r2 += std::max(lo * lo, hi * hi);
    }
    B.map_radius2 = float(r2 * (1.0 + 1e-4) + 1e-12);
  }
  size_t nv = size_t(res[0] + 1) * (res[1] + 1) * (res[2] + 1);
  size_t ne = size_t(res[0]) * (res[1] + 1) * (res[2] + 1) +
              size_t(res[0] + 1) * res[1] * (res[2] + 1) +
              size_t(res[0] + 1) * (res[1] + 1) * res[2];
  size_t nodes = nv + 2 * ne;
  const bool pellets = pile->volume_method == ALV_PELLETS;
  if (pellets) nodes = 0;  // build_grid builds no volume map (pile.hpp:523)
  if (nodes >= (1ull << 32)) return fail(ALV_ERR_UNSUPPORTED, "volume map too large");
  B.geom.num_nodes = unsigned(nodes);
  float* vol = nullptr;
  if (!pellets) {
    int rc = dev_alloc(&vol, nodes);
    if (rc != ALV_OK) return rc;
  }
  B.volume = vol;
  float* dmap = nullptr;
  if (dist->type == ALV_DISTANCE_MAP) {
    // construct_distance_grid (pile.hpp:79-95): the distance at every node
    int rc = dev_alloc(&dmap, nodes);
    if (rc != ALV_OK) {
      (void)hipFree(vol);
      return rc;
    }
    const unsigned chunk = 1u << 18;
    std::vector<float> pts(size_t(chunk) * 3), val(chunk);
    for (unsigned b0 = 0; b0 < B.geom.num_nodes && rc == ALV_OK; b0 += chunk) {
      const unsigned m = std::min(chunk, B.geom.num_nodes - b0);
      for (unsigned k = 0; k < m; ++k) {
        const vec3 xn = node_position(B.geom, b0 + k);
        pts[3 * size_t(k)] = xn.x;
        pts[3 * size_t(k) + 1] = xn.y;
        pts[3 * size_t(k) + 2] = xn.z;
      }
      dist->signed_distance(dist->context, pts.data(), m, val.data());
      if (hipMemcpy(dmap + b0, val.data(), size_t(m) * sizeof(float), hipMemcpyHostToDevice) !=
          hipSuccess)
        rc = fail(ALV_ERR_HIP, "distance map upload");
    }
    if (rc != ALV_OK) {
      (void)hipFree(vol);
      (void)hipFree(dmap);
      return rc;
    }
  }
  B.distance = dmap;
  for (int a = 0; a < 3; ++a) {
    B.x[a] = x ? x[a] : 0.f;
    B.v[a] = 0.f;
    B.omega[a] = 0.f;
  }
  for (int a = 0; a < 4; ++a) B.q[a] = q ? q[a] : (a == 3 ? 1.f : 0.f);
  B.mass = mass;
  B.dynamic = mass != 0.f ? 1 : 0;
  B.accum = B.dynamic;
  // build the volume map (update_volume_field, analytic variant)
  alv_consti cni;
  memset(&cni, 0, sizeof(cni));
  SimParams p = make_params(*cn, cni, 0, 0);
  const unsigned chunk = 1u << 20;
  for (unsigned b = 0; !pellets && b < B.geom.num_nodes; b += chunk) {
    unsigned e = std::min(B.geom.num_nodes, b + chunk);
    if (dmap)  // update_volume_field from the distance map (runner.hpp:1616-1658)
      launch_volume_field_grid(p, B.geom, dmap, vol, sign, pile->peirce, b, e, 0);
    else
      launch_volume_field_analytic(p, B, vol, pile->peirce, b, e, 0);
    ALV_HIP(hipGetLastError());
  }
  unsigned char* flags = nullptr;
  if (!pellets) {
    ALV_HIP(hipMalloc(&flags, std::max<size_t>(size_t(res[0]) * res[1] * res[2], 1)));
    launch_map_flags(B.geom, vol, flags, nullptr);
    ALV_HIP(hipGetLastError());
    ALV_HIP(hipDeviceSynchronize());
  }
  B.cell_nz = flags;
  std::array<float, 3> I = {inertia ? inertia[0] : 1.f, inertia ? inertia[1] : 1.f,
                            inertia ? inertia[2] : 1.f};
  if (slot < 0) {
    pile->bnd.push_back(B);
    pile->volumes.push_back(vol);
    pile->distances.push_back(dmap);
    pile->cell_nz.push_back(flags);
    pile->restitution.push_back(restitution);
    pile->friction.push_back(friction);
    pile->inertia.push_back(I);
    pile->force.push_back({0.f, 0.f, 0.f});
    pile->torque.push_back({0.f, 0.f, 0.f});
    pile->accel.push_back({0.f, 0.f, 0.f});
    slot = int(pile->bnd.size() - 1);
  } else {
    // the solvers read the boundary table at every step start, so the new
    // map and kinematics take effect from the next step
    ALV_HIP(hipDeviceSynchronize());
    if (pile->volumes[slot]) (void)hipFree(pile->volumes[slot]);
    if (pile->distances[slot]) (void)hipFree(pile->distances[slot]);
    if (pile->cell_nz[slot]) (void)hipFree(pile->cell_nz[slot]);
    pile->bnd[slot] = B;
    pile->volumes[slot] = vol;
    pile->distances[slot] = dmap;
    pile->cell_nz[slot] = flags;
    pile->restitution[slot] = restitution;
    pile->friction[slot] = friction;
    pile->inertia[slot] = I;
    pile->force[slot] = {0.f, 0.f, 0.f};
    pile->torque[slot] = {0.f, 0.f, 0.f};
    pile->accel[slot] = {0.f, 0.f, 0.f};
  }
  if (boundary_id) *boundary_id = unsigned(slot);
  return ALV_OK;
}
}  // namespace

extern "C" {

int alv_pile_add(alv_pile* pile, const alv_const* cn, const alv_distance* dist,
                 const unsigned resolution[3], float cell_width, float margin, float sign,
                 float mass, float restitution, float friction, const float inertia[3],
                 const float x[3], const float q[4], unsigned* boundary_id) {
  return place_boundary(pile, -1, cn, dist, resolution, cell_width, margin, sign, mass,
                        restitution, friction, inertia, x, q, boundary_id);
}

int alv_pile_replace(alv_pile* pile, unsigned b, const alv_const* cn, const alv_distance* dist,
                     const unsigned resolution[3], float cell_width, float margin, float sign,
                     float mass, float restitution, float friction, const float inertia[3],
                     const float x[3], const float q[4]) {
  return place_boundary(pile, int(b), cn, dist, resolution, cell_width, margin, sign, mass,
                        restitution, friction, inertia, x, q, nullptr);
}

int alv_pile_set_body(alv_pile* pile, unsigned b, float mass, float restitution, float friction,
                      const float inertia[3]) {
  if (!pile || b >= pile->bnd.size()) return fail(ALV_ERR_ARG, "bad boundary id");
  pile->bnd[b].mass = mass;
  pile->bnd[b].dynamic = mass != 0.f ? 1 : 0;
  pile->bnd[b].accum = pile->bnd[b].dynamic;
  pile->restitution[b] = restitution;
  pile->friction[b] = friction;
  if (inertia) pile->inertia[b] = {inertia[0], inertia[1], inertia[2]};
  return ALV_OK;
}

int alv_pile_get_body(const alv_pile* pile, unsigned b, float* mass, float* restitution,
                      float* friction, float inertia[3]) {
  if (!pile || b >= pile->bnd.size()) return fail(ALV_ERR_ARG, "bad boundary id");
  if (mass) *mass = pile->bnd[b].mass;
  if (restitution) *restitution = pile->restitution[b];
  if (friction) *friction = pile->friction[b];
  if (inertia)
    for (int a = 0; a < 3; ++a) inertia[a] = pile->inertia[b][a];
  return ALV_OK;
}

int alv_pile_compute_mask(alv_pile* pile, unsigned b, float distance_threshold,
                          const float* sample_x, float* mask, unsigned num_samples) {
  if (!pile || b >= pile->bnd.size()) return fail(ALV_ERR_ARG, "bad boundary id");
  if (num_samples && (!sample_x || !mask)) return fail(ALV_ERR_ARG, "null argument");
  // the host kinematics of the body, as Pile::compute_mask passes x_(i), q_(i)
  launch_boundary_mask(pile->bnd[b], sample_x, distance_threshold, mask, num_samples, nullptr);
  ALV_HIP(hipGetLastError());
  ALV_HIP(hipDeviceSynchronize());
  return ALV_OK;
}

int alv_pile_calculate_cfl_v2(const alv_pile* pile, float* max_v2) {
  if (!pile || !max_v2) return fail(ALV_ERR_ARG, "null argument");
  // Pile::calculate_cfl_v2 (pile.hpp:816-825): |omega x (max_distance, 0, 0) + v|^2
  float m = -FLT_MAX;
  for (const BoundaryDev& B : pile->bnd) {
    const float r[3] = {B.max_distance, 0.f, 0.f};
    const float c[3] = {B.omega[1] * r[2] - B.omega[2] * r[1] + B.v[0],
                        B.omega[2] * r[0] - B.omega[0] * r[2] + B.v[1],
                        B.omega[0] * r[1] - B.omega[1] * r[0] + B.v[2]};
    const float v2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
    if (v2 > m) m = v2;
  }
  *max_v2 = m;
  return ALV_OK;
}

int alv_pile_get_size(const alv_pile* pile, unsigned* size) {
  if (!pile || !size) return fail(ALV_ERR_ARG, "null argument");
  *size = unsigned(pile->bnd.size());
  return ALV_OK;
}

int alv_pile_set_kinematics(alv_pile* pile, unsigned b, const float x[3], const float v[3],
                            const float q[4], const float omega[3]) {
  if (!pile || b >= pile->bnd.size()) return fail(ALV_ERR_ARG, "bad boundary id");
  BoundaryDev& B = pile->bnd[b];
  for (int a = 0; a < 3; ++a) {
    if (x) B.x[a] = x[a];
    if (v) B.v[a] = v[a];
    if (omega) B.omega[a] = omega[a];
  }
  if (q)
    for (int a = 0; a < 4; ++a) B.q[a] = q[a];
  return ALV_OK;
}

int alv_pile_get_kinematics(const alv_pile* pile, unsigned b, float x[3], float v[3],
                            float q[4], float omega[3], float force[3], float torque[3]) {
  if (!pile || b >= pile->bnd.size()) return fail(ALV_ERR_ARG, "bad boundary id");
  const BoundaryDev& B = pile->bnd[b];
  for (int a = 0; a < 3; ++a) {
    if (x) x[a] = B.x[a];
    if (v) v[a] = B.v[a];
    if (omega) omega[a] = B.omega[a];
    if (force) force[a] = pile->force[b][a];
    if (torque) torque[a] = pile->torque[b][a];
  }
  if (q)
    for (int a = 0; a < 4; ++a) q[a] = B.q[a];
  return ALV_OK;
}

int alv_pile_set_gravity(alv_pile* pile, const float g[3]) {
  if (!pile || !g) return fail(ALV_ERR_ARG, "null argument");
  for (int a = 0; a < 3; ++a) pile->gravity[a] = g[a];
  return ALV_OK;
}

// Pile::integrate_kinematics (pile.hpp:794-815) with calculate_dq and
// calculate_angular_acceleration (runner.hpp:202-252).
int alv_pile_integrate_kinematics(alv_pile* pile, float dt) {
  if (!pile) return fail(ALV_ERR_ARG, "null pile");
  for (size_t b = 0; b < pile->bnd.size(); ++b) {
    BoundaryDev& B = pile->bnd[b];
    float4 q = make_float4(B.q[0], B.q[1], B.q[2], B.q[3]);
    auto dq_of = [](vec3 w, float4 qq) {
      return make_float4(0.5f * (w.x * qq.w + w.y * qq.z - w.z * qq.y),
                         0.5f * (-w.x * qq.z + w.y * qq.w + w.z * qq.x),
                         0.5f * (w.x * qq.y - w.y * qq.x + w.z * qq.w),
                         0.5f * (-w.x * qq.x - w.y * qq.y - w.z * qq.z));
    };
    auto qnorm = [](float4 qq) {
      float s = 1.0f / sqrtf(qq.x * qq.x + qq.y * qq.y + qq.z * qq.z + qq.w * qq.w);
      return make_float4(qq.x * s, qq.y * s, qq.z * s, qq.w * s);
    };
    vec3 xx = mk(B.x[0], B.x[1], B.x[2]);
    vec3 vv = mk(B.v[0], B.v[1], B.v[2]);
    vec3 ww = mk(B.omega[0], B.omega[1], B.omega[2]);
    if (B.mass == 0.f) {
      xx += vv * dt;
      float4 dq = dq_of(ww, q);
      q = make_float4(q.x + dt * dq.x, q.y + dt * dq.y, q.z + dt * dq.z, q.w + dt * dq.w);
      q = qnorm(q);
    } else {
      vec3 F = mk(pile->force[b][0], pile->force[b][1], pile->force[b][2]);
      vec3 T = mk(pile->torque[b][0], pile->torque[b][1], pile->torque[b][2]);
      vv += ((1.0f / B.mass) * F) * dt;
      vec3 local = rotate(T, conjugate(q));
      const auto& I = pile->inertia[b];
      vec3 scaled = mk(local.x / I[0], local.y / I[1], local.z / I[2]);
      ww += rotate(scaled, q) * dt;
      vec3 acc = mk(pile->gravity[0], pile->gravity[1], pile->gravity[2]);
      pile->accel[b] = {acc.x, acc.y, acc.z};
      float4 dq = dq_of(ww, q);
      q = make_float4(q.x + dt * dq.x, q.y + dt * dq.y, q.z + dt * dq.z, q.w + dt * dq.w);
      q = qnorm(q);
      vec3 dx = (acc * dt + vv) * dt;
      xx += dx;
      vv = (1.0f / dt) * dx;
    }
    B.x[0] = xx.x;
    B.x[1] = xx.y;
    B.x[2] = xx.z;
    B.v[0] = vv.x;
    B.v[1] = vv.y;
    B.v[2] = vv.z;
    B.omega[0] = ww.x;
    B.omega[1] = ww.y;
    B.omega[2] = ww.z;
    B.q[0] = q.x;
    B.q[1] = q.y;
    B.q[2] = q.z;
    B.q[3] = q.w;
  }
  return ALV_OK;
}

}  // extern "C"

namespace {
// Grows a device buffer to hold at least `count` elements of T.
template <class T>
hipError_t reserve_dev(T** ptr, size_t* cap, size_t count) {
  if (*ptr && *cap >= count) return hipSuccess;
  if (*ptr) (void)hipFree(*ptr);
  *ptr = nullptr;
  *cap = 0;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(ptr), std::max<size_t>(count, 1) * sizeof(T));
  if (e == hipSuccess) *cap = count;
  return e;
}

// Uploads the collision vertices (body frame, concatenated) when they
// changed; vbase[b] is body b's first vertex.
int upload_vertices(alv_pile* P, std::vector<unsigned>& vbase, hipStream_t stream) {
  const unsigned nb = unsigned(P->bnd.size());
  P->coll_verts.resize(nb);
  vbase.assign(nb + 1, 0);
  for (unsigned b = 0; b < nb; ++b) vbase[b + 1] = vbase[b] + unsigned(P->coll_verts[b].size() / 3);
  if (P->verts_dirty) {
    std::vector<float> all;
    all.reserve(size_t(vbase[nb]) * 3);
    std::vector<unsigned> body;
    body.reserve(vbase[nb]);
    for (unsigned b = 0; b < nb; ++b) {
      all.insert(all.end(), P->coll_verts[b].begin(), P->coll_verts[b].end());
      body.insert(body.end(), vbase[b + 1] - vbase[b], b);
    }
    ALV_HIP(reserve_dev(&P->verts_dev, &P->verts_cap, all.size()));
    ALV_HIP(reserve_dev(&P->vbody_dev, &P->vbody_cap, body.size()));
    if (!all.empty()) {
      ALV_HIP(hipMemcpyAsync(P->verts_dev, all.data(), all.size() * sizeof(float),
                             hipMemcpyHostToDevice, stream));
      ALV_HIP(hipMemcpyAsync(P->vbody_dev, body.data(), body.size() * sizeof(unsigned),
                             hipMemcpyHostToDevice, stream));
    }
    ALV_HIP(hipStreamSynchronize(stream));
    P->verts_dirty = false;
  }
  return ALV_OK;
}

// Runs the contact kernel over `jobs` with the body kinematics of `bodies`
// and leaves the kept contacts in P->contacts.
int run_collision(alv_pile* P, const std::vector<ContactJob>& jobs, unsigned candidates,
                  const BoundaryDev* bodies, const float* verts, float tolerance,
                  hipStream_t stream) {
  const unsigned nb = unsigned(P->bnd.size());
  P->num_found = 0;
  P->contacts.clear();
  if (jobs.empty()) return ALV_OK;
  ALV_HIP(reserve_dev(&P->bodies_dev, &P->bodies_cap, nb));
  ALV_HIP(reserve_dev(&P->inertia_dev, &P->inertia_cap, size_t(nb) * 5));
  ALV_HIP(reserve_dev(&P->jobs_dev, &P->jobs_cap, jobs.size()));
  if (!P->count_dev) ALV_HIP(hipMalloc(&P->count_dev, sizeof(unsigned)));
  if (!P->contacts_dev && P->max_num_contacts)
    ALV_HIP(hipMalloc(&P->contacts_dev, size_t(P->max_num_contacts) * sizeof(alv_contact)));
  std::vector<float> inertia(size_t(nb) * 5);
  for (unsigned b = 0; b < nb; ++b) {
    for (int k = 0; k < 3; ++k) inertia[b * 3 + k] = P->inertia[b][k];
    inertia[nb * 3 + b * 2] = P->restitution[b];
    inertia[nb * 3 + b * 2 + 1] = P->friction[b];
  }
  ALV_HIP(hipMemcpyAsync(P->bodies_dev, bodies, nb * sizeof(BoundaryDev), hipMemcpyHostToDevice,
                         stream));
  ALV_HIP(hipMemcpyAsync(P->inertia_dev, inertia.data(), inertia.size() * sizeof(float),
                         hipMemcpyHostToDevice, stream));
  ALV_HIP(hipMemcpyAsync(P->jobs_dev, jobs.data(), jobs.size() * sizeof(ContactJob),
                         hipMemcpyHostToDevice, stream));
  launch_collision(P->jobs_dev, unsigned(jobs.size()), candidates, nb, P->bodies_dev,
                   P->inertia_dev, verts, tolerance, P->max_num_contacts, P->contacts_dev,
                   P->count_dev, stream);
  ALV_HIP(hipGetLastError());
  unsigned found = 0;
  ALV_HIP(hipMemcpyAsync(&found, P->count_dev, sizeof(unsigned), hipMemcpyDeviceToHost, stream));
  ALV_HIP(hipStreamSynchronize(stream));
  P->num_found = found;
  const unsigned kept = std::min(found, P->max_num_contacts);
  P->contacts.resize(kept);
  if (kept) {
    ALV_HIP(hipMemcpyAsync(P->contacts.data(), P->contacts_dev, kept * sizeof(alv_contact),
                           hipMemcpyDeviceToHost, stream));
    ALV_HIP(hipStreamSynchronize(stream));
  }
  return ALV_OK;
}

// Pile::find_contacts (pile.hpp:826-904) on `stream`; see alv.h for the
// meaning of i and j.  Leaves the kept contacts in pile->contacts.
int find_contacts(alv_pile* P, int i, int j, float tolerance, hipStream_t stream) {
  const unsigned nb = unsigned(P->bnd.size());
  if (i >= int(nb) || j >= int(nb)) return fail(ALV_ERR_ARG, "bad boundary id");
  std::vector<unsigned> vbase;
  int rc = upload_vertices(P, vbase, stream);
  if (rc != ALV_OK) return rc;
  std::vector<ContactJob> jobs;
  unsigned candidates = 0;
  const unsigned i0 = i < 0 ? 0 : unsigned(i), i1 = i < 0 ? nb : unsigned(i) + 1;
  for (unsigned a = i0; a < i1; ++a) {
    const unsigned j0 = j < 0 ? 0 : unsigned(j), j1 = j < 0 ? nb : unsigned(j) + 1;
    for (unsigned b = j0; b < j1; ++b) {
      if (j < 0 && (a == b || (P->bnd[a].mass == 0.f && P->bnd[b].mass == 0.f))) continue;
      if (P->bnd[b].type == ALV_DISTANCE_BOX_SHELL) continue;
      const unsigned nv = vbase[a + 1] - vbase[a];
      if (nv == 0) continue;
      ContactJob J{};
      J.i = a;
      J.j = b;
      J.vbase = vbase[a];
      J.nv = nv;
      J.cbase = candidates;
      J.nj = 1;
      jobs.push_back(J);
      candidates += nv;
    }
  }
  return run_collision(P, jobs, candidates, P->bnd.data(), P->verts_dev, tolerance, stream);
}

// Pile::find_contacts(particle_x, max_num_beads), pellet branch
// (pile.hpp:858-886; collision_test_with_pellets, runner.hpp:3627-3727):
// the world pellets (`world`, pellet-major, as placed at the step start)
// against every body j, or against each hinted run of identical bodies in
// one job, with the kinematics of the last device copy.  Contacts come out
// in (run, pellet, j) order.  A run uses each body's own distance, which is
// the run's first body's when the hint holds.
int find_contacts_pellets(alv_pile* P, const float* world, float tolerance, hipStream_t stream) {
  const unsigned nb = unsigned(P->bnd.size());
  std::vector<unsigned> vbase;
  int rc = upload_vertices(P, vbase, stream);
  if (rc != ALV_OK) return rc;
  if (P->kin_snap.size() != size_t(nb) * 13) P->snapshot_kinematics();
  std::vector<BoundaryDev> bodies(P->bnd);
  for (unsigned b = 0; b < nb; ++b)
    for (int a = 0; a < 4; ++a) {
      if (a < 3) {
        bodies[b].x[a] = P->kin_snap[b * 3 + a];
        bodies[b].v[a] = P->kin_snap[nb * 3 + b * 3 + a];
        bodies[b].omega[a] = P->kin_snap[nb * 10 + b * 3 + a];
      }
      bodies[b].q[a] = P->kin_snap[nb * 6 + b * 4 + a];
    }
  std::vector<ContactJob> jobs;
  unsigned candidates = 0;
  unsigned jn = 0, seq = 0;
  while (jn < nb) {
    unsigned jb = jn, je = jn + 1;
    if (seq < P->ident.size() && jn == P->ident[seq].first && jn < P->ident[seq].second) {
      jb = P->ident[seq].first;
      je = std::min(nb, P->ident[seq].second);
      ++seq;
    }
    if (P->bnd[jb].type != ALV_DISTANCE_BOX_SHELL) {
      for (unsigned a = 0; a < nb; ++a) {
        const unsigned nv = vbase[a + 1] - vbase[a];
        if (nv == 0) continue;
        ContactJob J{};
        J.i = a;
        J.j = jb;
        J.vbase = vbase[a];
        J.nv = nv;
        J.cbase = candidates;
        J.nj = je - jb;
        J.flags = kJobWorld | kJobSkip;
        jobs.push_back(J);
        candidates += nv * (je - jb);
      }
    }
    jn = je;
  }
  return run_collision(P, jobs, candidates, bodies.data(), world, tolerance, stream);
}

// Pile::solve_contacts (pile.hpp:906-979): sequential impulse sweeps with a
// penetration correction and dynamic friction, on the host kinematics.
void solve_contacts(alv_pile* P) {
  if (P->num_found > P->max_num_contacts)
    fprintf(stderr, "No. of contacts exceeds %u.\n", P->max_num_contacts);
  auto v3 = [](const float* a) { return mk(a[0], a[1], a[2]); };
  auto add3 = [](float* a, vec3 d) {
    a[0] += d.x;
    a[1] += d.y;
    a[2] += d.z;
  };
  auto sym = [](vec3 v, vec3 d, vec3 o) {
    return mk(v.x * d.x + v.y * o.x + v.z * o.y, v.x * o.x + v.y * d.y + v.z * o.z,
              v.x * o.y + v.y * o.z + v.z * d.z);
  };
  for (int sweep = 0; sweep < 5; ++sweep) {
    for (alv_contact& C : P->contacts) {
      BoundaryDev& Bi = P->bnd[C.i];
      BoundaryDev& Bj = P->bnd[C.j];
      const vec3 n = v3(C.n), t = v3(C.t), cp_i = v3(C.cp_i), cp_j = v3(C.cp_j);
      const float depth = dot(n, cp_i - cp_j);
      const vec3 r_i = cp_i - v3(Bi.x), r_j = cp_j - v3(Bj.x);
      const vec3 u_rel = (v3(Bi.v) + cross(v3(Bi.omega), r_i)) - (v3(Bj.v) + cross(v3(Bj.omega), r_j));
      const float u_rel_n = dot(n, u_rel);
      float corr = C.nkninv * (C.goalu - u_rel_n);
      if (corr < -C.impulse_sum) corr = -C.impulse_sum;
      if (depth < 0) corr -= C.nkninv * depth;  // unit stiffness
      vec3 p = corr * n;
      C.impulse_sum += corr;
      const float pn = dot(p, n);
      if (C.friction * pn > C.pmax)
        p -= C.pmax * t;
      else if (C.friction * pn < -C.pmax)
        p += C.pmax * t;
      else
        p -= (C.friction * pn) * t;
      if (Bi.mass != 0.f) {
        add3(Bi.v, p * (1.0f / Bi.mass));
        add3(Bi.omega, sym(cross(r_i, p), v3(C.iiwi_diag), v3(C.iiwi_off_diag)));
      }
      if (Bj.mass != 0.f) {
        add3(Bj.v, -p * (1.0f / Bj.mass));
        add3(Bj.omega, sym(cross(r_j, -p), v3(C.iiwj_diag), v3(C.iiwj_off_diag)));
      }
    }
  }
}

unsigned pellet_count(const alv_pile* P) {
  size_t n = 0;
  for (const auto& v : P->coll_verts) n += v.size() / 3;
  return unsigned(n);
}

bool has_collision_vertices(const alv_pile* P) {
  for (const auto& v : P->coll_verts)
    if (!v.empty()) return true;
  return false;
}
}  // namespace

extern "C" {

int alv_pile_set_collision_vertices(alv_pile* pile, unsigned b, const float* xyz, unsigned n) {
  if (!pile || b >= pile->bnd.size()) return fail(ALV_ERR_ARG, "bad boundary id");
  if (n && !xyz) return fail(ALV_ERR_ARG, "null vertices");
  pile->coll_verts.resize(pile->bnd.size());
  if (pile->volume_method == ALV_PELLETS) {
    size_t total = n;
    for (size_t k = 0; k < pile->coll_verts.size(); ++k)
      if (k != b) total += pile->coll_verts[k].size() / 3;
    if (total > pile->max_pellets)  // pile.hpp:266-275
      return fail(ALV_ERR_ARG, "Cannot add pellets because the new pellet count " +
                                   std::to_string(total) + " exceeds the limit " +
                                   std::to_string(pile->max_pellets));
  }
  pile->coll_verts[b].assign(xyz, xyz + size_t(n) * 3);
  pile->verts_dirty = true;
  return ALV_OK;
}

int alv_pile_num_pellets(const alv_pile* pile, unsigned* num_pellets, unsigned* max_num_pellets) {
  if (!pile) return fail(ALV_ERR_ARG, "null pile");
  if (num_pellets) *num_pellets = pile->volume_method == ALV_PELLETS ? pellet_count(pile) : 0u;
  if (max_num_pellets) *max_num_pellets = pile->max_pellets;
  return ALV_OK;
}

int alv_pile_hint_identical_sequence(alv_pile* pile, unsigned begin, unsigned end) {
  if (!pile) return fail(ALV_ERR_ARG, "null pile");
  pile->ident.push_back({begin, end});
  return ALV_OK;
}

int alv_pile_find_contacts_pellets(alv_pile* pile, const float* particle_x,
                                   unsigned max_num_beads, float contact_tolerance,
                                   unsigned* num_contacts) {
  if (!pile) return fail(ALV_ERR_ARG, "null pile");
  int rc;
  if (pile->volume_method == ALV_PELLETS) {
    if (!particle_x) return fail(ALV_ERR_ARG, "null particle_x");
    rc = find_contacts_pellets(pile, particle_x + size_t(max_num_beads) * 3, contact_tolerance,
                               nullptr);
  } else {
    rc = find_contacts(pile, -1, -1, contact_tolerance, nullptr);
  }
  if (rc == ALV_OK && num_contacts) *num_contacts = pile->num_found;
  return rc;
}

int alv_pile_find_contacts(alv_pile* pile, int i, int j, float contact_tolerance,
                           unsigned* num_contacts) {
  if (!pile) return fail(ALV_ERR_ARG, "null pile");
  if (i < 0 && j >= 0) return fail(ALV_ERR_ARG, "j given without i");
  int rc = find_contacts(pile, i, j, contact_tolerance, nullptr);
  if (rc == ALV_OK && num_contacts) *num_contacts = pile->num_found;
  return rc;
}

int alv_pile_solve_contacts(alv_pile* pile) {
  if (!pile) return fail(ALV_ERR_ARG, "null pile");
  solve_contacts(pile);
  return ALV_OK;
}

int alv_pile_copy_kinematics_to_device(alv_pile* pile) {
  if (!pile) return fail(ALV_ERR_ARG, "null pile");
  pile->snapshot_kinematics();
  float *x, *v, *q, *w;
  return alv_pile_kinematics_device(pile, &x, &v, &q, &w);
}

int alv_pile_kinematics_device(alv_pile* pile, float** x, float** v, float** q, float** omega) {
  if (!pile || !x || !v || !q || !omega) return fail(ALV_ERR_ARG, "null argument");
  const size_t B = pile->bnd.size();
  if (pile->kin_snap.size() != B * 13) pile->snapshot_kinematics();
  ALV_HIP(reserve_dev(&pile->kin_dev, &pile->kin_cap, std::max<size_t>(B * 13, 1)));
  if (pile->kin_dirty && B) {
    ALV_HIP(hipMemcpy(pile->kin_dev, pile->kin_snap.data(), B * 13 * sizeof(float),
                      hipMemcpyHostToDevice));
    pile->kin_dirty = false;
  }
  *x = pile->kin_dev;
  *v = pile->kin_dev + B * 3;
  *q = pile->kin_dev + B * 6;
  *omega = pile->kin_dev + B * 10;
  return ALV_OK;
}

int alv_pile_get_contacts(const alv_pile* pile, alv_contact* out, unsigned capacity, unsigned* n) {
  if (!pile || (capacity && !out)) return fail(ALV_ERR_ARG, "null argument");
  const unsigned k = std::min(capacity, unsigned(pile->contacts.size()));
  if (k) memcpy(out, pile->contacts.data(), k * sizeof(alv_contact));
  if (n) *n = k;
  return ALV_OK;
}

int alv_pile_get_volume_grid(const alv_pile* pile, unsigned b, float** nodes,
                             unsigned* num_nodes) {
  if (!pile || b >= pile->bnd.size()) return fail(ALV_ERR_ARG, "bad boundary id");
  if (nodes) {
    *nodes = pile->volumes[b];
    if (pile->volumes[b]) pile->maps_dirty = true;  // the caller may write the nodes
  }
  if (num_nodes) *num_nodes = pile->bnd[b].geom.num_nodes;
  return ALV_OK;
}

int alv_pile_get_distance_grid(const alv_pile* pile, unsigned b, float** nodes,
                               unsigned* num_nodes) {
  if (!pile || b >= pile->bnd.size()) return fail(ALV_ERR_ARG, "bad boundary id");
  const bool map = pile->distances[b] != nullptr;
  if (nodes) *nodes = pile->distances[b];
  if (num_nodes) *num_nodes = map ? pile->bnd[b].geom.num_nodes : 0u;
  return ALV_OK;
}

int alv_pile_get_grid_info(const alv_pile* pile, unsigned b, float dmin[3], float dmax[3],
                           unsigned res[3], float cs[3], float* sign, float* mass) {
  if (!pile || b >= pile->bnd.size()) return fail(ALV_ERR_ARG, "bad boundary id");
  const BoundaryDev& B = pile->bnd[b];
  for (int a = 0; a < 3; ++a) {
    if (dmin) dmin[a] = B.geom.dmin[a];
    if (dmax) dmax[a] = B.geom.dmax[a];
    if (res) res[a] = B.geom.res[a];
    if (cs) cs[a] = B.geom.cs[a];
  }
  if (sign) *sign = B.sign;
  if (mass) *mass = B.mass;
  return ALV_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Solver
// ---------------------------------------------------------------------------
static int solver_alloc(alv_solver* S) {
  const unsigned cap = S->cap;
  const unsigned C = S->num_cells;
  const unsigned K = S->cni.max_num_neighbors_per_particle;
  DevArrays& A = S->A;
  memset(&A, 0, sizeof(A));
  A.cap = cap;
  // positions, velocities and pressure hold the pellets after the fluid
  // (solver.hpp:30-36)
  const unsigned pcap = S->pcap;
  const size_t UC = size_t(cap) + pcap;
  int rc = ALV_OK;
#define ALV_ALLOC(ptr, cnt) \
  if ((rc = S->alloc(&(ptr), (cnt))) != ALV_OK) return rc;
  ALV_ALLOC(A.ux, UC * 3);
  ALV_ALLOC(A.uv, UC * 3);
  ALV_ALLOC(A.ua, size_t(cap) * 3);
  ALV_ALLOC(A.udensity, cap);
  ALV_ALLOC(A.ualpha, cap);
  ALV_ALLOC(A.ukappa, cap);
  ALV_ALLOC(A.ukappa_v, cap);
  ALV_ALLOC(A.urho_adv, cap);
  ALV_ALLOC(A.unnb, cap);
  ALV_ALLOC(A.upressure, UC);
  ALV_ALLOC(A.key, cap);
  ALV_ALLOC(A.slot, cap);
  ALV_ALLOC(A.sorted_id, cap);
  ALV_ALLOC(A.sort_tmp, cap);
  const size_t CT = S->cells_total;
  ALV_ALLOC(A.cell_count, CT);
  ALV_ALLOC(A.cell_start, CT + 1);
  ALV_ALLOC(A.scan_tmp, (CT + kScanPerBlock - 1) / kScanPerBlock + 1);
  const size_t NS = S->scenes.size();
  ALV_ALLOC(S->par_dev, NS);
  ALV_ALLOC(S->blk_scene, size_t(grid_for(cap)) + 1);
  ALV_ALLOC(S->cblk_scene, CT / kBlock + 1);
  ALV_ALLOC(A.anomaly, 4);
  ALV_ALLOC(A.x, cap);
  ALV_ALLOC(A.v, cap);
  ALV_ALLOC(A.a, cap);
  ALV_ALLOC(A.density, cap);
  ALV_ALLOC(A.alpha, cap);
  ALV_ALLOC(A.kappa, cap);
  ALV_ALLOC(A.kappa_v, cap);
  ALV_ALLOC(A.rho_adv, cap);
  ALV_ALLOC(A.kfac, cap);
  ALV_ALLOC(A.feat, size_t(cap) * 2);
  ALV_ALLOC(A.tiles, size_t(grid_for(cap)) + 1);
  A.use_tiles = 0;  // LDS tiles are opt-in (alv_solver_set_tiles); the autotuner tries them
  ALV_ALLOC(A.kpred, cap);
  A.kmax = K;
  ALV_ALLOC(A.nbr, nbr_capacity(K, cap));
  ALV_ALLOC(A.nnb, cap);
  ALV_ALLOC(A.bmask, cap);
  size_t max_bnd = 0, all_bnd = 0;
  for (const auto& sc : S->scenes) {
    max_bnd = std::max(max_bnd, sc.pile->bnd.size());
    all_bnd += sc.pile->bnd.size();
  }
  const size_t BN = std::max<size_t>(1, max_bnd) * cap;
  ALV_ALLOC(A.brec, BN);
  ALV_ALLOC(A.bker, BN);
  ALV_ALLOC(A.bforce, BN);
  ALV_ALLOC(A.btorque, BN);
  if (S->kind != ALV_SOLVER_DF) {  // IISPH and ISPH pressure solves
    ALV_ALLOC(A.pressure, cap);
    ALV_ALLOC(A.last_pressure, cap);
    ALV_ALLOC(A.aii, cap);
    ALV_ALLOC(A.dii, cap);
    ALV_ALLOC(A.dij_pj, cap);
    ALV_ALLOC(A.adv_density, cap);
    ALV_ALLOC(A.density_err, cap);
    ALV_ALLOC(A.pacc, cap);
  }
  ALV_ALLOC(A.partials, grid_for(cap) + 1);
  ALV_ALLOC(A.fpartials, grid_for(cap) + 1);
  ALV_ALLOC(A.body_partials, std::max<size_t>(1, max_bnd) * 256 * 6);
  if (pcap) {
    A.pcap = pcap;
    const size_t pspan = size_t(grid_for(pcap)) * kBlock;
    ALV_ALLOC(A.pkey, pcap);
    ALV_ALLOC(A.pslot, pcap);
    ALV_ALLOC(A.psorted_id, pspan);
    ALV_ALLOC(A.psort_tmp, pspan);
    ALV_ALLOC(A.pcell_count, CT);
    ALV_ALLOC(A.pcell_start, CT + 1);
    ALV_ALLOC(A.pscan_tmp, (CT + kScanPerBlock - 1) / kScanPerBlock + 1);
    ALV_ALLOC(A.px, pspan);
    ALV_ALLOC(A.pv, pspan);
    ALV_ALLOC(A.pbody, pspan);
    ALV_ALLOC(A.pp, pspan);
    ALV_ALLOC(A.pnbr, nbr_capacity(K, cap));
    ALV_ALLOC(A.pnnb, cap);
    ALV_ALLOC(A.pbnbr, nbr_capacity(K, pcap));
    ALV_ALLOC(A.pbnnb, pspan);
    ALV_ALLOC(S->ppar_dev, 1);
    ALV_ALLOC(S->pblk_scene, size_t(grid_for(pcap)) + 1);
    ALV_HIP(hipMemset(S->pblk_scene, 0, (size_t(grid_for(pcap)) + 1) * sizeof(unsigned short)));
  }
#undef ALV_ALLOC
  S->bnd_capacity = unsigned(std::max<size_t>(1, all_bnd));
  // scene of every particle block and every cell block
  {
    std::vector<unsigned short> bs(size_t(grid_for(cap)) + 1, 0), cs(CT / kBlock + 1, 0);
    for (size_t k = 0; k < NS; ++k) {
      const auto& sc = S->scenes[k];
      for (unsigned b = sc.first / kBlock; b < (sc.first + sc.span) / kBlock; ++b)
        bs[b] = (unsigned short)k;
      const unsigned cspan = (k + 1 < NS ? S->scenes[k + 1].cell_base : unsigned(CT)) - sc.cell_base;
      for (unsigned b = sc.cell_base / kBlock; b < (sc.cell_base + cspan) / kBlock; ++b)
        cs[b] = (unsigned short)k;
    }
    ALV_HIP(hipMemcpy(S->blk_scene, bs.data(), bs.size() * sizeof(unsigned short),
                      hipMemcpyHostToDevice));
    ALV_HIP(hipMemcpy(S->cblk_scene, cs.data(), cs.size() * sizeof(unsigned short),
                      hipMemcpyHostToDevice));
  }
  ALV_HIP(hipHostMalloc(reinterpret_cast<void**>(&S->par_host), NS * sizeof(SimParams),
                        hipHostMallocDefault));
  memset(S->par_host, 0, NS * sizeof(SimParams));
  // The control blocks and body sums come back through coherent pinned
  // memory that the kernels store into (no copy command per readback).
  static_assert(sizeof(StepCtl) % alignof(BoundaryDev) == 0, "boundaries follow the blocks");
  const size_t step_bytes = NS * sizeof(StepCtl) + S->bnd_capacity * sizeof(BoundaryDev);
  ALV_HIP(hipMalloc(&S->step_dev, step_bytes));
  ALV_HIP(hipHostMalloc(&S->step_host, step_bytes, hipHostMallocCoherent));
  memset(S->step_host, 0, step_bytes);
  S->ctl_host = static_cast<StepCtl*>(S->step_host);
  S->bnd_host = reinterpret_cast<BoundaryDev*>(S->ctl_host + NS);
  A.ctl = static_cast<StepCtl*>(S->step_dev);
  S->bnd_dev = reinterpret_cast<BoundaryDev*>(A.ctl + NS);
  ALV_HIP(hipHostMalloc(reinterpret_cast<void**>(&S->body_host),
                        S->bnd_capacity * 6 * sizeof(double), hipHostMallocCoherent));
  memset(S->body_host, 0, S->bnd_capacity * 6 * sizeof(double));
  ALV_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&A.ctl_pub), S->ctl_host, 0));
  ALV_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&A.body_sums), S->body_host, 0));
  A.bnd = S->bnd_dev;
  return ALV_OK;
}

extern "C" {

int alv_solver_create(alv_solver** out, int kind, alv_pile* pile, unsigned max_num_particles,
                      const alv_const* cn, const alv_consti* cni) {
  if (!out || !pile || !cn || !cni) return fail(ALV_ERR_ARG, "null argument");
  if (kind != ALV_SOLVER_DF && kind != ALV_SOLVER_II && kind != ALV_SOLVER_I)
    return fail(ALV_ERR_ARG, "unknown solver kind");
  if (max_num_particles == 0) return fail(ALV_ERR_ARG, "max_num_particles is zero");
  if (cni->max_num_neighbors_per_particle == 0)
    return fail(ALV_ERR_ARG, "max_num_neighbors_per_particle is zero");
  size_t C = size_t(cni->grid_res[0]) * cni->grid_res[1] * cni->grid_res[2];
  if (C == 0 || C > (1ull << 31)) return fail(ALV_ERR_ARG, "grid_res out of range");
  if (pile->volume_method == ALV_PELLETS && kind != ALV_SOLVER_I)
    return fail(ALV_ERR_UNSUPPORTED, "pellet boundaries step with the ISPH solver (SolverI) only");
  alv_solver* S = new (std::nothrow) alv_solver();
  if (!S) return fail(ALV_ERR_STATE, "out of host memory");
  S->kind = kind;
  S->pile = pile;
  S->cap = max_num_particles;
  S->pcap = pile->max_pellets;
  S->cn = *cn;
  S->cni = *cni;
  S->num_cells = unsigned(C);
  {
    alv_solver::Scene sc;
    memset(&sc, 0, sizeof(sc));
    sc.pile = pile;
    sc.n = max_num_particles;
    sc.first = 0;
    sc.span = grid_for(max_num_particles) * kBlock;
    sc.cell_base = 0;
    sc.num_cells = unsigned(C);
    for (int a = 0; a < 3; ++a) {
      sc.off[a] = cni->grid_offset[a];
      sc.res[a] = cni->grid_res[a];
    }
    S->scenes.push_back(sc);
    S->cells_total = grid_for(unsigned(C) + 1) * kBlock;
  }
  memset(&S->set, 0, sizeof(S->set));
  S->set.num_particles = max_num_particles;
  S->set.density_change_tolerance = 1e-3f;
  S->set.density_error_tolerance = 1e-3f;
  S->set.min_density_solve = 2;
  S->set.max_density_solve = 100;
  S->set.min_divergence_solve = 1;
  S->set.max_divergence_solve = 100;
  S->set.enable_divergence_solve = 1;
  S->set.enable_density_solve = 1;
  memset(&S->rep, 0, sizeof(S->rep));
  hipError_t e = hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete S;
    return fail(ALV_ERR_HIP, hipGetErrorString(e));
  }
  int rc = solver_alloc(S);
  if (rc != ALV_OK) {
    delete S;
    return rc;
  }
  *out = S;
  return ALV_OK;
}

int alv_solver_destroy(alv_solver* solver) {
  delete solver;
  return ALV_OK;
}

int alv_solver_create_batch(alv_solver** out, int kind, unsigned num_scenes,
                            const alv_scene_desc* descs, const alv_const* cn,
                            const alv_consti* cni) {
  if (!out || !descs || !cn || !cni) return fail(ALV_ERR_ARG, "null argument");
  if (kind != ALV_SOLVER_DF && kind != ALV_SOLVER_II && kind != ALV_SOLVER_I)
    return fail(ALV_ERR_ARG, "unknown solver kind");
  if (num_scenes == 0 || num_scenes > 65535) return fail(ALV_ERR_ARG, "num_scenes out of range");
  if (cni->max_num_neighbors_per_particle == 0)
    return fail(ALV_ERR_ARG, "max_num_neighbors_per_particle is zero");
  std::unique_ptr<alv_solver> S(new (std::nothrow) alv_solver());
  if (!S) return fail(ALV_ERR_STATE, "out of host memory");
  size_t first = 0, cells = 0, bnd = 0;
  for (unsigned k = 0; k < num_scenes; ++k) {
    const alv_scene_desc& d = descs[k];
    if (!d.pile) return fail(ALV_ERR_ARG, "scene without a pile");
    if (d.pile->volume_method == ALV_PELLETS)
      return fail(ALV_ERR_UNSUPPORTED, "batched scenes use volume-map boundaries");
    if (d.num_particles == 0) return fail(ALV_ERR_ARG, "scene without particles");
    for (const auto& B : d.pile->bnd)
      if (B.dynamic)
        return fail(ALV_ERR_UNSUPPORTED, "batched scenes take static boundaries only");
    const size_t C = size_t(d.grid_res[0]) * d.grid_res[1] * d.grid_res[2];
    if (C == 0) return fail(ALV_ERR_ARG, "grid_res out of range");
    alv_solver::Scene sc;
    memset(&sc, 0, sizeof(sc));
    sc.pile = d.pile;
    sc.n = d.num_particles;
    sc.first = unsigned(first);
    sc.span = grid_for(d.num_particles) * kBlock;
    sc.cell_base = unsigned(cells);
    sc.num_cells = unsigned(C);
    for (int a = 0; a < 3; ++a) {
      sc.off[a] = d.grid_offset[a];
      sc.res[a] = d.grid_res[a];
    }
    sc.bnd_base = unsigned(bnd);
    sc.last_div = 2;
    sc.last_dens = 3;
    first += sc.span;
    cells += size_t(grid_for(unsigned(C) + 1)) * kBlock;
    bnd += d.pile->bnd.size();
    if (first > (1ull << 31) || cells > (1ull << 31))
      return fail(ALV_ERR_UNSUPPORTED, "batch too large");
    S->scenes.push_back(sc);
  }
  S->batch = true;
  S->kind = kind;
  S->pile = descs[0].pile;
  S->cap = unsigned(first);
  S->cn = *cn;
  S->cni = *cni;
  for (int a = 0; a < 3; ++a) {
    S->cni.grid_res[a] = descs[0].grid_res[a];
    S->cni.grid_offset[a] = descs[0].grid_offset[a];
  }
  S->num_cells = S->scenes[0].num_cells;
  S->cells_total = unsigned(cells);
  memset(&S->set, 0, sizeof(S->set));
  S->set.num_particles = S->cap;
  S->set.density_change_tolerance = 1e-3f;
  S->set.density_error_tolerance = 1e-3f;
  S->set.min_density_solve = 2;
  S->set.max_density_solve = 100;
  S->set.min_divergence_solve = 1;
  S->set.max_divergence_solve = 100;
  S->set.enable_divergence_solve = 1;
  S->set.enable_density_solve = 1;
  memset(&S->rep, 0, sizeof(S->rep));
  hipError_t e = hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return fail(ALV_ERR_HIP, hipGetErrorString(e));
  int rc = solver_alloc(S.get());
  if (rc != ALV_OK) return rc;
  *out = S.release();
  return ALV_OK;
}

int alv_solver_num_scenes(const alv_solver* S, unsigned* n) {
  if (!S || !n) return fail(ALV_ERR_ARG, "null argument");
  *n = unsigned(S->scenes.size());
  return ALV_OK;
}

int alv_solver_scene_range(const alv_solver* S, unsigned scene, unsigned* first,
                           unsigned* count) {
  if (!S) return fail(ALV_ERR_ARG, "null solver");
  if (scene >= S->scenes.size()) return fail(ALV_ERR_ARG, "scene out of range");
  const auto& sc = S->scenes[scene];
  if (first) *first = sc.first;
  if (count) *count = S->batch ? sc.n : S->set.num_particles;
  return ALV_OK;
}

int alv_solver_scene_report(const alv_solver* S, unsigned scene, alv_solver_report* r) {
  if (!S || !r) return fail(ALV_ERR_ARG, "null argument");
  if (scene >= S->scenes.size()) return fail(ALV_ERR_ARG, "scene out of range");
  if (!S->batch) return alv_solver_get_report(S, r);
  *r = S->scenes[scene].rep;
  return ALV_OK;
}

int alv_solver_set_const(alv_solver* S, const alv_const* cn, const alv_consti* cni) {
  if (!S || !cn) return fail(ALV_ERR_ARG, "null argument");
  S->cn = *cn;
  if (cni && S->batch) {
    // every scene keeps the grid it was created with
    if (cni->max_num_neighbors_per_particle != S->cni.max_num_neighbors_per_particle)
      return fail(ALV_ERR_STATE, "max_num_neighbors_per_particle is fixed at solver construction");
    S->cni.max_num_particles_per_cell = cni->max_num_particles_per_cell;
  } else if (cni) {
    if (cni->grid_res[0] != S->cni.grid_res[0] || cni->grid_res[1] != S->cni.grid_res[1] ||
        cni->grid_res[2] != S->cni.grid_res[2] ||
        cni->max_num_neighbors_per_particle != S->cni.max_num_neighbors_per_particle)
      return fail(ALV_ERR_STATE, "grid_res and max_num_neighbors_per_particle are fixed at "
                                 "solver construction");
    S->cni = *cni;
  }
  return ALV_OK;
}

int alv_solver_set_settings(alv_solver* S, const alv_solver_settings* st) {
  if (!S || !st) return fail(ALV_ERR_ARG, "null argument");
  if (st->num_particles > S->cap)
    return fail(ALV_ERR_ARG, "num_particles exceeds max_num_particles");
  if (S->batch) {
    // a batch keeps every scene's particles; a new dt or t applies to all
    // scenes
    for (auto& sc : S->scenes) {
      if (st->dt != S->set.dt || sc.dt == 0.0f) sc.dt = st->dt;
      if (st->t != S->set.t) sc.t = st->t;
    }
    S->set = *st;
    S->set.num_particles = S->cap;
    return ALV_OK;
  }
  S->set = *st;
  return ALV_OK;
}

int alv_solver_get_settings(const alv_solver* S, alv_solver_settings* st) {
  if (!S || !st) return fail(ALV_ERR_ARG, "null argument");
  *st = S->set;
  return ALV_OK;
}

int alv_solver_get_report(const alv_solver* S, alv_solver_report* r) {
  if (!S || !r) return fail(ALV_ERR_ARG, "null argument");
  *r = S->rep;
  r->t = S->set.t;
  r->dt = S->set.dt;
  return ALV_OK;
}

}  // extern "C"

namespace {
// Scene parameters of every scene, uploaded when they changed, and the
// launch description.  A single-scene solver follows num_particles.
int solver_launch(alv_solver* S, Launch* L) {
  if (!S->batch) {
    auto& sc = S->scenes[0];
    sc.n = S->set.num_particles;
    sc.span = grid_for(sc.n) * kBlock;
  }
  const size_t ns = S->scenes.size();
  bool changed = !S->par_valid;
  unsigned pblocks = 0, active = 0;
  for (size_t k = 0; k < ns; ++k) {
    const auto& sc = S->scenes[k];
    SimParams p = make_params(S->cn, S->cni, unsigned(sc.pile->bnd.size()), sc.n);
    for (int a = 0; a < 3; ++a) {
      p.res[a] = int(sc.res[a]);
      p.off[a] = sc.off[a];
    }
    p.num_cells = sc.num_cells;
    p.first = sc.first;
    p.span = sc.span;
    p.cell_base = sc.cell_base;
    p.bnd_base = sc.bnd_base;
    if (sc.pile->volume_method == ALV_PELLETS) {
      p.pellet_mode = 1;
      p.np = pellet_count(sc.pile);
      if (p.np > S->pcap)
        return fail(ALV_ERR_STATE, "pellets were added beyond max_num_pellets of the solver");
    }
    if (memcmp(&p, &S->par_host[k], sizeof(p)) != 0) {
      S->par_host[k] = p;
      changed = true;
    }
    pblocks = std::max(pblocks, (sc.first + sc.span) / kBlock);
    active += sc.n;
  }
  if (changed) {
    ALV_HIP(hipMemcpyAsync(S->par_dev, S->par_host, ns * sizeof(SimParams),
                           hipMemcpyHostToDevice, S->stream));
    S->par_valid = 1;
  }
  L->T = SceneTab{S->par_dev, S->blk_scene, S->cblk_scene};
  L->pblocks = pblocks;
  L->cells = S->cells_total;
  L->cblocks = S->cells_total / kBlock;
  L->scenes = unsigned(ns);
  L->active = active;
  L->tiles = S->A.use_tiles;
  L->p0 = S->par_host[0];
  return ALV_OK;
}

int upload_step_inputs(alv_solver* S) {
  size_t nb = 0;
  for (const auto& sc : S->scenes) {
    if (sc.pile->maps_dirty) {
      for (size_t b = 0; b < sc.pile->bnd.size(); ++b)
        launch_map_flags(sc.pile->bnd[b].geom, sc.pile->volumes[b], sc.pile->cell_nz[b], S->stream);
      ALV_HIP(hipGetLastError());
      sc.pile->maps_dirty = false;
    }
  }
  for (const auto& sc : S->scenes) {
    const size_t b = sc.pile->bnd.size();
    if (sc.bnd_base + b > S->bnd_capacity)
      return fail(ALV_ERR_STATE, "boundaries were added after the solver was created");
    if (b) memcpy(S->bnd_host + sc.bnd_base, sc.pile->bnd.data(), b * sizeof(BoundaryDev));
    for (size_t k = 0; k < b; ++k) {
      BoundaryDev& B = S->bnd_host[sc.bnd_base + k];
      B.accum = B.dynamic || S->record_static_forces;
    }
    nb = std::max<size_t>(nb, sc.bnd_base + b);
  }
  for (size_t k = 0; k < S->scenes.size(); ++k) {
    StepCtl* c = S->ctl_host + k;
    memset(c, 0, sizeof(StepCtl));
    c->dt = S->batch ? S->scenes[k].dt : S->set.dt;
    c->cfl = S->set.cfl;
    c->min_dt = S->set.min_dt;
    c->max_dt = S->set.max_dt;
    c->tol_div = S->set.density_change_tolerance;
    c->tol_dens = S->set.density_error_tolerance;
    c->min_div = S->set.min_divergence_solve;
    c->max_div = S->set.max_divergence_solve;
    c->min_dens = S->set.min_density_solve;
    c->max_dens = S->set.max_density_solve;
    c->length_scale = S->cn.particle_radius * 2.0f;
  }
  // the control blocks and the boundaries in one copy
  ALV_HIP(hipMemcpyAsync(S->step_dev, S->step_host,
                         S->scenes.size() * sizeof(StepCtl) + nb * sizeof(BoundaryDev),
                         hipMemcpyHostToDevice, S->stream));
  return ALV_OK;
}

// Every kernel that changes a control block publishes it to ctl_host
// (publish_ctl), so the blocks are current once the stream has drained.
int read_ctl(alv_solver* S) {
  ALV_HIP(hipStreamSynchronize(S->stream));
  return ALV_OK;
}

bool all_done(const alv_solver* S, bool divergence) {
  for (size_t k = 0; k < S->scenes.size(); ++k) {
    const StepCtl& c = S->ctl_host[k];
    if (!(divergence ? c.div_done : c.dens_done)) return false;
  }
  return true;
}

// Launch set of the pellets of a pellet solver (one scene): the fluid's
// constants with n = the pellet count; uploads the pile's pellets.
int pellet_launch(alv_solver* S, const Launch& L, Launch* Lp) {
  std::vector<unsigned> vbase;
  int rc = upload_vertices(S->pile, vbase, S->stream);
  if (rc != ALV_OK) return rc;
  S->A.pverts = S->pile->verts_dev;
  S->A.pvbody = S->pile->vbody_dev;
  SimParams pp = L.p0;
  pp.n = pp.np;
  pp.first = 0;
  pp.span = grid_for(pp.np) * kBlock;
  ALV_HIP(hipMemcpyAsync(S->ppar_dev, &pp, sizeof(pp), hipMemcpyHostToDevice, S->stream));
  // the copy reads pp before this function returns
  ALV_HIP(hipStreamSynchronize(S->stream));
  *Lp = L;
  Lp->T = SceneTab{S->ppar_dev, S->pblk_scene, S->cblk_scene};
  Lp->pblocks = grid_for(pp.np);
  Lp->active = pp.np;
  Lp->p0 = pp;
  return ALV_OK;
}

// grid + gather + boundary records + neighbour list + density (+ factor).
// Only kernels (no memset nodes), so the sequence can be captured in a graph.
// gather_state: also gather the solver state the step's passes read in
// sorted order (stiffnesses or pressure); a neighbour update on its own
// leaves it out, since only a step reads it, after its own neighbour phase.
int neighbor_phase(alv_solver* S, const Launch& L, int with_factor, bool gather_state = true) {
  DevArrays& A = S->A;
  launch_clear_grid(L, A, S->stream);
  if (L.active == 0) return ALV_OK;
  launch_cell_count(L, A, S->stream);
  launch_reserve_padding(L, A, S->stream);
  launch_scan(L.cells, A.cell_count, A.cell_start, A.scan_tmp, S->stream);
  launch_scatter(L, A, S->stream);
  launch_cell_sort(L, A, S->stream);
  launch_tiles(L, A, S->stream);
  const int kind = !gather_state            ? ALV_KIND_NONE
                   : S->kind != ALV_SOLVER_DF ? ALV_KIND_II
                   : with_factor             ? ALV_KIND_DF
                                             : ALV_KIND_DF_FACTOR;
  launch_gather_boundary(L, A, kind, S->stream);
  launch_neighbors(L, A, with_factor, S->stream);
  ALV_HIP(hipGetLastError());
  S->grid_valid = 1;
  return ALV_OK;
}

// Enqueues Jacobi iterations in chunks until every scene's device flag says
// done.
template <class LaunchFn>
int run_loop(alv_solver* S, bool divergence, unsigned max_iter, unsigned* last,
             LaunchFn launch) {
  unsigned issued = 0;
  unsigned chunk = std::max(1u, *last);
  while (issued < max_iter) {
    unsigned todo = std::min(chunk, max_iter - issued);
    for (unsigned k = 0; k < todo; ++k) launch();
    ALV_HIP(hipGetLastError());
    issued += todo;
    int rc = read_ctl(S);
    if (rc != ALV_OK) return rc;
    if (all_done(S, divergence)) break;
    chunk = std::max(1u, chunk / 2);
  }
  return ALV_OK;
}

// A DFSPH step with a solve switched off leaves that solve's count and mean
// as they were (the reference sets them inside the solve's enable test,
// solver_df.hpp:129-130, 294-295).
void fill_report(const alv_solver* S, alv_solver_report* r, const StepCtl& c) {
  const bool df = S->kind == ALV_SOLVER_DF;
  if (!df || S->set.enable_divergence_solve) {
    r->num_divergence_solve = c.div_iter;
    r->mean_density_change = c.mean_density_change;
  }
  if (!df || S->set.enable_density_solve) {
    r->num_density_solve = c.dens_iter;
    r->mean_density_error = c.mean_density_error;
  }
  r->particle_max_v2 = c.particle_max_v2;
  r->pile_max_v2 = c.pile_max_v2;
  r->max_v2 = c.max_v2;
  r->cfl_dt = c.cfl_dt;
  r->utilized_cfl = c.utilized_cfl;
}

int finish_step(alv_solver* S, const Launch& L) {
  // rigid force sums (solver_df.hpp:347-353): k_body_sums stores them in
  // body_host, so the control blocks' wait brings back both
  alv_pile* pile = S->pile;
  bool any_dynamic = false;
  for (const auto& B : pile->bnd) any_dynamic |= B.dynamic != 0;
  const bool body_sums = !S->batch && any_dynamic && L.active;
  if (body_sums) {
    launch_body_sums(L, S->A, S->stream);
    ALV_HIP(hipGetLastError());
  }
  int rc = read_ctl(S);
  if (rc != ALV_OK) return rc;
  if (g_event_profiler) g_event_profiler->collect();
  S->dirty = ALV_F_DENSITY | ALV_F_ALPHA | ALV_F_RHO_ADV | ALV_F_A | ALV_F_NNB;
  S->export_stale = ~0u;
  if (S->batch) {
    // every scene keeps its own step, time and report; the solver's report
    // is the maximum iteration counts over the scenes
    unsigned mdiv = 0, mdens = 0;
    for (size_t k = 0; k < S->scenes.size(); ++k) {
      auto& sc = S->scenes[k];
      const StepCtl& c = S->ctl_host[k];
      fill_report(S, &sc.rep, c);
      sc.rep.t = sc.t + c.dt;
      sc.rep.dt = c.dt;
      sc.t += c.dt;
      sc.dt = c.dt;
      mdiv = std::max(mdiv, c.div_iter);
      mdens = std::max(mdens, c.dens_iter);
    }
    fill_report(S, &S->rep, S->ctl_host[0]);
    const bool df = S->kind == ALV_SOLVER_DF;
    if (!df || S->set.enable_divergence_solve) S->rep.num_divergence_solve = mdiv;
    if (!df || S->set.enable_density_solve) S->rep.num_density_solve = mdens;
    if (mdiv) S->last_div = mdiv;
    if (mdens) S->last_dens = mdens;
    S->set.dt = S->scenes[0].dt;
    S->set.t = S->scenes[0].t;
    return ALV_OK;
  }
  const StepCtl& c = *S->ctl_host;
  S->set.dt = c.dt;
  fill_report(S, &S->rep, c);
  if (c.div_iter) S->last_div = c.div_iter;
  if (c.dens_iter) S->last_dens = c.dens_iter;
  if (body_sums) {
    for (size_t b = 0; b < pile->bnd.size(); ++b) {
      if (!pile->bnd[b].dynamic) continue;
      for (int a = 0; a < 3; ++a) {
        pile->force[b][a] = float(S->body_host[b * 6 + a]);
        pile->torque[b][a] = float(S->body_host[b * 6 + 3 + a]);
      }
    }
  }
  alv_pile_integrate_kinematics(pile, S->set.dt);
  // rigid contacts (solver_df.hpp:356-357, solver_ii.hpp:240-241)
  if (has_collision_vertices(pile)) {
    if (pile->volume_method == ALV_PELLETS)  // pellets as placed at the step start
      rc = find_contacts_pellets(pile, S->A.ux + size_t(S->cap) * 3, S->cn.contact_tolerance,
                                 S->stream);
    else
      rc = find_contacts(pile, -1, -1, S->cn.contact_tolerance, S->stream);
    if (rc != ALV_OK) return rc;
    solve_contacts(pile);
  }
  S->set.t += S->set.dt;
  return ALV_OK;
}
}  // namespace

extern "C" {

}  // extern "C"

namespace {
// Runs one segment of a step: enqueued directly, or replayed from a graph
// captured from the same enqueue sequence (profiling needs the direct path).
template <class Enqueue>
int run_segment(alv_solver* S, int which, unsigned long long key, Enqueue enqueue) {
  if (!S->use_graph || g_profiler) {
    enqueue();
    ALV_HIP(hipGetLastError());
    return ALV_OK;
  }
  alv_solver::Segment& G = S->seg[which];
  if (!G.exec || G.key != key) {
    if (G.exec) {
      ALV_HIP(hipGraphExecDestroy(G.exec));
      G.exec = nullptr;
    }
    hipGraph_t graph = nullptr;
    ALV_HIP(hipStreamBeginCapture(S->stream, hipStreamCaptureModeThreadLocal));
    enqueue();
    hipError_t e = hipStreamEndCapture(S->stream, &graph);
    if (e != hipSuccess) return fail(ALV_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&G.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      G.exec = nullptr;
      return fail(ALV_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
    G.key = key;
  }
  ALV_HIP(hipGraphLaunch(G.exec, S->stream));
  return ALV_OK;
}

// Iterations enqueued before the first convergence check: the last step's
// count plus the slack; grows at once, shrinks after 16 smaller steps (so a
// captured graph is not rebuilt for every small change).
unsigned next_chunk(unsigned last, unsigned slack, unsigned lo, unsigned hi, unsigned* chunk,
                    unsigned* shrink) {
  const unsigned want = std::min(hi, std::max(std::max(lo, 1u), last + slack));
  if (want > *chunk || *chunk == 0 || *chunk > hi) {
    *chunk = want;
    *shrink = 0;
  } else if (want < *chunk && ++*shrink >= 16) {
    *chunk = want;
    *shrink = 0;
  }
  return *chunk;
}

// Continues a Jacobi loop after its first chunk until every scene is done.
template <class LaunchFn>
int finish_loop(alv_solver* S, bool divergence, unsigned issued, unsigned max_iter,
                LaunchFn launch) {
  int rc = read_ctl(S);
  if (rc != ALV_OK) return rc;
  unsigned chunk = 1;
  while (!all_done(S, divergence) && issued < max_iter) {
    const unsigned todo = std::min(chunk, max_iter - issued);
    for (unsigned k = 0; k < todo; ++k) launch();
    ALV_HIP(hipGetLastError());
    issued += todo;
    rc = read_ctl(S);
    if (rc != ALV_OK) return rc;
    chunk = std::min(chunk * 2, 8u);
  }
  return ALV_OK;
}

// Staleness bits of the reference-layout copies (alv_solver::export_stale).
enum : unsigned { kXNbr = 1u, kXBnd = 2u, kXIi = 4u };

// Reference-layout arrays made on request (ALV_FIELD_NEIGHBORS ..
// ALV_FIELD_PRESSURE_ACCEL): allocated on first use, rebuilt when the state
// moved on since the last copy.
int export_field(alv_solver* S, int field, void** dptr, size_t* n_elements,
                 unsigned* n_components) {
  const size_t cap = S->cap;
  const size_t K = S->cni.max_num_neighbors_per_particle;
  size_t nbnd = 1;
  for (const auto& sc : S->scenes) nbnd = std::max<size_t>(nbnd, sc.pile->bnd.size());
  const bool ii_field = field >= ALV_FIELD_LAST_PRESSURE;
  if (ii_field && S->kind != ALV_SOLVER_II)
    return fail(ALV_ERR_ARG, "the IISPH arrays belong to SolverIi");
  unsigned bit = field == ALV_FIELD_NEIGHBORS ? kXNbr : ii_field ? kXIi : kXBnd;
  if (bit == kXNbr && !S->export_nbr) ALV_HIP(hipMalloc(&S->export_nbr, cap * K * sizeof(float4)));
  if (bit == kXBnd && !S->export_brec) {
    ALV_HIP(hipMalloc(&S->export_brec, nbnd * cap * sizeof(float4)));
    ALV_HIP(hipMalloc(&S->export_bker, nbnd * cap * sizeof(float4)));
    ALV_HIP(hipMalloc(&S->export_force, nbnd * cap * 3 * sizeof(float)));
    ALV_HIP(hipMalloc(&S->export_torque, nbnd * cap * 3 * sizeof(float)));
  }
  if (bit == kXIi && !S->export_ii_scalar[0]) {
    for (float*& p : S->export_ii_scalar) ALV_HIP(hipMalloc(&p, cap * sizeof(float)));
    for (float*& p : S->export_ii_vec) ALV_HIP(hipMalloc(&p, cap * 3 * sizeof(float)));
  }
  if (S->export_stale & bit) {
    // rows of unused slots stay zero
    if (bit == kXNbr) ALV_HIP(hipMemsetAsync(S->export_nbr, 0, cap * K * sizeof(float4), S->stream));
    if (bit == kXBnd) {
      ALV_HIP(hipMemsetAsync(S->export_brec, 0, nbnd * cap * sizeof(float4), S->stream));
      ALV_HIP(hipMemsetAsync(S->export_bker, 0, nbnd * cap * sizeof(float4), S->stream));
      ALV_HIP(hipMemsetAsync(S->export_force, 0, nbnd * cap * 3 * sizeof(float), S->stream));
      ALV_HIP(hipMemsetAsync(S->export_torque, 0, nbnd * cap * 3 * sizeof(float), S->stream));
    }
    if (bit == kXIi) {
      for (float* p : S->export_ii_scalar)
        ALV_HIP(hipMemsetAsync(p, 0, cap * sizeof(float), S->stream));
      for (float* p : S->export_ii_vec)
        ALV_HIP(hipMemsetAsync(p, 0, cap * 3 * sizeof(float), S->stream));
    }
    Launch L;
    int rc = solver_launch(S, &L);
    if (rc != ALV_OK) return rc;
    if (L.active && S->grid_valid) {
      if (bit == kXNbr) launch_export_neighbor_rows(L, S->A, S->export_nbr, S->stream);
      if (bit == kXBnd)
        launch_export_boundary(L, S->A, S->export_brec, S->export_bker, S->export_force,
                               S->export_torque, S->stream);
      if (bit == kXIi) launch_export_iisph(L, S->A, S->export_ii_scalar, S->export_ii_vec, S->stream);
      ALV_HIP(hipGetLastError());
    }
    ALV_HIP(hipStreamSynchronize(S->stream));
    S->export_stale &= ~bit;
  }
  void* ptr = nullptr;
  size_t n = cap;
  unsigned comp = 1;
  switch (field) {
    case ALV_FIELD_NEIGHBORS: ptr = S->export_nbr; n = cap * K; comp = 4; break;
    case ALV_FIELD_BOUNDARY: ptr = S->export_brec; n = nbnd * cap; comp = 4; break;
    case ALV_FIELD_BOUNDARY_KERNEL: ptr = S->export_bker; n = nbnd * cap; comp = 4; break;
    case ALV_FIELD_FORCE: ptr = S->export_force; n = nbnd * cap; comp = 3; break;
    case ALV_FIELD_TORQUE: ptr = S->export_torque; n = nbnd * cap; comp = 3; break;
    case ALV_FIELD_LAST_PRESSURE: ptr = S->export_ii_scalar[0]; break;
    case ALV_FIELD_AII: ptr = S->export_ii_scalar[1]; break;
    case ALV_FIELD_ADV_DENSITY: ptr = S->export_ii_scalar[2]; break;
    case ALV_FIELD_DENSITY_ERR: ptr = S->export_ii_scalar[3]; break;
    case ALV_FIELD_DII: ptr = S->export_ii_vec[0]; comp = 3; break;
    case ALV_FIELD_DIJ_PJ: ptr = S->export_ii_vec[1]; comp = 3; break;
    case ALV_FIELD_PRESSURE_ACCEL: ptr = S->export_ii_vec[2]; comp = 3; break;
    default: return fail(ALV_ERR_ARG, "unknown field");
  }
  *dptr = ptr;
  if (n_elements) *n_elements = n;
  if (n_components) *n_components = comp;
  return ALV_OK;
}
}  // namespace

extern "C" {

int alv_solver_set_static_boundary_forces(alv_solver* S, int enable) {
  if (!S) return fail(ALV_ERR_ARG, "null solver");
  S->record_static_forces = enable ? 1 : 0;
  return ALV_OK;
}

int alv_solver_step(alv_solver* S) {
  if (!S) return fail(ALV_ERR_ARG, "null solver");
  Launch L;
  int rc = solver_launch(S, &L);
  if (rc != ALV_OK) return rc;
  rc = upload_step_inputs(S);
  if (rc != ALV_OK) return rc;
  DevArrays& A = S->A;
  hipStream_t st = S->stream;
  if (!S->batch) S->pile->snapshot_kinematics();  // Pile::copy_kinematics_to_device
  if (S->kind != ALV_SOLVER_DF && (!A.pressure || !A.last_pressure || !A.aii || !A.pacc))
    return fail(ALV_ERR_STATE, "pressure-solve arrays are not allocated");
  if (S->kind == ALV_SOLVER_DF) {
    // SolverDf::step (solver_df.hpp:74-359) in three segments; the host
    // checks convergence between them
    const int div = S->set.enable_divergence_solve != 0;
    const int dens = S->set.enable_density_solve != 0;
    const int learned_dens = S->net != nullptr && S->net_dens && div;
    const unsigned cdiv = next_chunk(S->last_div, S->slack, S->set.min_divergence_solve,
                                     S->set.max_divergence_solve, &S->chunk_div, &S->shrink_div);
    const unsigned cdens = next_chunk(S->last_dens, S->slack, S->set.min_density_solve,
                                      S->set.max_density_solve, &S->chunk_dens, &S->shrink_dens);
    const MlpDev mv = mlp_device_view(S->net);
    const unsigned long long base = (unsigned long long)L.pblocks << 24 ^
                                    (unsigned long long)(L.active != 0) << 23 ^
                                    (unsigned long long)(L.p0.wrap_mask != 0) << 22;
    const unsigned long long keyA =
        base ^ ((unsigned long long)cdiv << 8) ^ (div << 1) ^ (S->log_features << 2) ^
        ((S->net != nullptr) << 3) ^ (S->net_div << 4) ^
        ((unsigned long long)(uintptr_t)S->net << 32);
    // the reference's warm start 0 rides on the neighbour search's walk over
    // the lists unless the learned warm start or feature logging replaces it
    const int fuse_ws = div && !S->net && !S->log_features;
    // Without the divergence solve the reference does not compute the DFSPH
    // factor (solver_df.hpp:81-91): the density solve reads the one each
    // particle ended the last step with.  It goes to user order with the
    // last step's permutation here and comes back in the new one
    // (k_gather_boundary, ALV_KIND_DF_FACTOR).
    if (!div && L.active && (S->dirty & ALV_F_ALPHA)) {
      launch_unsort_fields(L, A, ALV_F_ALPHA, st);
      ALV_HIP(hipGetLastError());
      S->dirty &= ~ALV_F_ALPHA;
    }
    rc = run_segment(S, 0, keyA, [&] {
      if (neighbor_phase(S, L, fuse_ws ? 2 : div) != ALV_OK) return;
      if (L.active && div) {
        launch_df_warm_start_div(L, A, st, S->log_features, S->net ? &mv : nullptr, S->net_div,
                                 fuse_ws);
        for (unsigned k = 0; k < cdiv; ++k) launch_df_div_iteration(L, A, st);
      }
    });
    if (rc != ALV_OK) return rc;
    if (L.active && div) {
      rc = finish_loop(S, true, cdiv, S->set.max_divergence_solve,
                       [&] { launch_df_div_iteration(L, A, st); });
      if (rc != ALV_OK) return rc;
    }
    const unsigned long long keyB = base ^ ((unsigned long long)cdens << 8) ^ (div << 1) ^
                                    (dens << 2) ^ (learned_dens << 3);
    rc = run_segment(S, 1, keyB, [&] {
      // with the divergence solve, its density-change passes computed the
      // viscosity
      if (L.active && !div) launch_df_viscosity(L, A, st);
      launch_update_dt(L, A, st);
      if (L.active) {
        launch_df_integrate_np(L, A, div, st);
        if (dens) {
          launch_df_warm_start_dens(L, A, st, learned_dens);
          for (unsigned k = 0; k < cdens; ++k) launch_df_dens_iteration(L, A, st);
        }
      }
    });
    if (rc != ALV_OK) return rc;
    if (L.active && dens) {
      rc = finish_loop(S, false, cdens, S->set.max_density_solve,
                       [&] { launch_df_dens_iteration(L, A, st); });
      if (rc != ALV_OK) return rc;
    }
    rc = run_segment(S, 2, base ^ (dens << 1), [&] {
      if (L.active) launch_df_finish(L, A, dens, st);
    });
    if (rc != ALV_OK) return rc;
  } else if (S->kind == ALV_SOLVER_I && L.p0.pellet_mode) {
    // SolverI::step (solver_i.hpp:79-330) with pellets
    Launch Lp;
    rc = pellet_launch(S, L, &Lp);
    if (rc != ALV_OK) return rc;
    launch_transform_pellets(Lp, A, st);  // transform_all_pellets
    launch_pellet_grid(Lp, A, st);
    rc = neighbor_phase(S, L, 0);
    if (rc != ALV_OK) return rc;
    launch_pellet_lists(L, Lp, A, st);
    if (L.active) launch_pellet_viscosity(L, A, st);
    launch_update_dt(L, A, st);
    if (L.active) {
      launch_is_prepare_pellets(L, A, st);
      rc = run_loop(S, false, S->set.max_density_solve, &S->last_dens,
                    [&] { launch_is_iteration_pellets(L, Lp, A, st); });
      if (rc != ALV_OK) return rc;
      launch_is_settle(L, A, st);
      launch_is_finish_pellets(L, Lp, A, st);
    }
  } else if (S->kind == ALV_SOLVER_I) {
    // SolverI::step (solver_i.hpp:79-330), volume-map boundaries
    rc = neighbor_phase(S, L, 0);
    if (rc != ALV_OK) return rc;
    if (L.active) launch_ii_viscosity_predict(L, A, st);
    launch_update_dt(L, A, st);
    if (L.active) {
      launch_is_prepare(L, A, st);
      rc = run_loop(S, false, S->set.max_density_solve, &S->last_dens,
                    [&] { launch_is_iteration(L, A, st); });
      if (rc != ALV_OK) return rc;
      launch_is_settle(L, A, st);
      launch_ii_finish(L, A, st);
    }
  } else {
    // SolverIi::step (solver_ii.hpp:76-243)
    rc = neighbor_phase(S, L, 0);
    if (rc != ALV_OK) return rc;
    if (L.active) launch_ii_viscosity_predict(L, A, st);
    launch_update_dt(L, A, st);
    if (L.active) {
      launch_ii_predict(L, A, st);
      rc = run_loop(S, false, S->set.max_density_solve, &S->last_dens,
                    [&] { launch_ii_iteration(L, A, st); });
      if (rc != ALV_OK) return rc;
      launch_ii_finish(L, A, st);
    }
  }
  ALV_HIP(hipGetLastError());
  S->features_valid = S->kind == ALV_SOLVER_DF && S->set.enable_divergence_solve &&
                      (S->log_features || S->net) && L.active;
  return finish_step(S, L);
}

int alv_solver_set_graph(alv_solver* S, int enable, unsigned slack) {
  if (!S) return fail(ALV_ERR_ARG, "null solver");
  S->use_graph = enable != 0;
  S->slack = slack;
  return ALV_OK;
}

int alv_solver_set_tiles(alv_solver* S, int enable) {
  if (!S) return fail(ALV_ERR_ARG, "null solver");
  ALV_HIP(hipStreamSynchronize(S->stream));
  S->A.use_tiles = enable != 0;
  for (auto& g : S->seg)  // captured graphs hold the old flag
    g.key = ~0ull;
  return ALV_OK;
}

int alv_solver_set_warm_start_net(alv_solver* S, alv_mlp* net, int use_divergence,
                                  int use_density) {
  if (!S) return fail(ALV_ERR_ARG, "null solver");
  if (net && S->kind != ALV_SOLVER_DF)
    return fail(ALV_ERR_UNSUPPORTED, "the learned warm start is a DFSPH feature");
  S->net_div = net && use_divergence;
  S->net_dens = net && use_density;
  S->net = (S->net_div || S->net_dens) ? net : nullptr;
  return ALV_OK;
}

int alv_solver_set_feature_log(alv_solver* S, int enable) {
  if (!S) return fail(ALV_ERR_ARG, "null solver");
  if (enable && S->kind != ALV_SOLVER_DF)
    return fail(ALV_ERR_UNSUPPORTED, "warm-start features are a DFSPH feature");
  S->log_features = enable != 0;
  return ALV_OK;
}

int alv_solver_export_training_pairs(alv_solver* S, float* dev_x, float* dev_y) {
  if (!S || !dev_x || !dev_y) return fail(ALV_ERR_ARG, "null argument");
  if (!S->features_valid)
    return fail(ALV_ERR_STATE,
                "no features recorded: enable the feature log (or a warm-start network) "
                "and take a step with the divergence solve enabled");
  Launch L;
  int rc = solver_launch(S, &L);
  if (rc != ALV_OK) return rc;
  launch_export_pairs(L, S->A, dev_x, dev_y, S->stream);
  ALV_HIP(hipGetLastError());
  ALV_HIP(hipStreamSynchronize(S->stream));
  return ALV_OK;
}

int alv_solver_update_particle_neighbors(alv_solver* S) {
  if (!S) return fail(ALV_ERR_ARG, "null solver");
  Launch L;
  int rc = solver_launch(S, &L);
  if (rc != ALV_OK) return rc;
  rc = upload_step_inputs(S);
  if (rc != ALV_OK) return rc;
  Launch Lp;
  if (L.p0.pellet_mode) {
    // the pellets at their current particle_x rows (solver.hpp:229-259)
    rc = pellet_launch(S, L, &Lp);
    if (rc != ALV_OK) return rc;
    launch_pellet_grid(Lp, S->A, S->stream);
  }
  rc = neighbor_phase(S, L, S->kind == ALV_SOLVER_DF ? 1 : 0, false);
  if (rc != ALV_OK) return rc;
  if (L.p0.pellet_mode) launch_pellet_lists(L, Lp, S->A, S->stream);
  ALV_HIP(hipGetLastError());
  ALV_HIP(hipStreamSynchronize(S->stream));
  S->dirty |= ALV_F_DENSITY | ALV_F_NNB | ALV_F_ALPHA;
  S->export_stale = ~0u;
  return ALV_OK;
}

int alv_solver_transform_all_pellets(alv_solver* S) {
  if (!S) return fail(ALV_ERR_ARG, "null solver");
  Launch L;
  int rc = solver_launch(S, &L);
  if (rc != ALV_OK) return rc;
  if (!L.p0.pellet_mode) return ALV_OK;
  rc = upload_step_inputs(S);
  if (rc != ALV_OK) return rc;
  Launch Lp;
  rc = pellet_launch(S, L, &Lp);
  if (rc != ALV_OK) return rc;
  launch_transform_pellets(Lp, S->A, S->stream);
  ALV_HIP(hipGetLastError());
  ALV_HIP(hipStreamSynchronize(S->stream));
  return ALV_OK;
}

int alv_solver_compute_density_change(alv_solver* S) {
  if (!S) return fail(ALV_ERR_ARG, "null solver");
  Launch L;
  int rc = solver_launch(S, &L);
  if (rc != ALV_OK) return rc;
  rc = upload_step_inputs(S);
  if (rc != ALV_OK) return rc;
  if (L.active) launch_density_change(L, S->A, S->stream);
  ALV_HIP(hipGetLastError());
  ALV_HIP(hipStreamSynchronize(S->stream));
  S->dirty |= ALV_F_RHO_ADV;
  return ALV_OK;
}

int alv_solver_integrate_velocity(alv_solver* S) {
  if (!S) return fail(ALV_ERR_ARG, "null solver");
  Launch L;
  int rc = solver_launch(S, &L);
  if (rc != ALV_OK) return rc;
  if (L.active) launch_integrate_velocity(L, S->A, S->set.dt, S->stream);
  ALV_HIP(hipGetLastError());
  ALV_HIP(hipStreamSynchronize(S->stream));
  return ALV_OK;
}

int alv_solver_reset_solving_var(alv_solver* S) {
  if (!S) return fail(ALV_ERR_ARG, "null solver");
  ALV_HIP(hipStreamSynchronize(S->stream));
  if (S->kind == ALV_SOLVER_DF) {
    ALV_HIP(hipMemset(S->A.ukappa, 0, size_t(S->cap) * sizeof(float)));
    ALV_HIP(hipMemset(S->A.ukappa_v, 0, size_t(S->cap) * sizeof(float)));
  } else {
    ALV_HIP(hipMemset(S->A.upressure, 0, size_t(S->cap) * sizeof(float)));
  }
  return ALV_OK;
}

int alv_solver_reset_t(alv_solver* S) {
  if (!S) return fail(ALV_ERR_ARG, "null solver");
  S->set.t = 0.f;
  S->set.dt = S->set.initial_dt;
  return ALV_OK;
}

int alv_solver_get_field(alv_solver* S, int field, void** dptr, size_t* n_elements,
                         unsigned* n_components) {
  if (!S || !dptr) return fail(ALV_ERR_ARG, "null argument");
  DevArrays& A = S->A;
  const size_t N = S->cap;
  unsigned need = 0;
  void* ptr = nullptr;
  size_t n = N;
  unsigned comp = 1;
  switch (field) {
    case ALV_FIELD_X: ptr = A.ux; comp = 3; n = N + S->pcap; break;
    case ALV_FIELD_V: ptr = A.uv; comp = 3; n = N + S->pcap; break;
    case ALV_FIELD_A: ptr = A.ua; comp = 3; need = ALV_F_A; break;
    case ALV_FIELD_DENSITY: ptr = A.udensity; need = ALV_F_DENSITY; break;
    case ALV_FIELD_DFSPH_FACTOR: ptr = A.ualpha; need = ALV_F_ALPHA; break;
    case ALV_FIELD_KAPPA: ptr = A.ukappa; break;
    case ALV_FIELD_KAPPA_V: ptr = A.ukappa_v; break;
    case ALV_FIELD_DENSITY_ADV: ptr = A.urho_adv; need = ALV_F_RHO_ADV; break;
    case ALV_FIELD_NUM_NEIGHBORS: ptr = A.unnb; need = ALV_F_NNB; break;
    case ALV_FIELD_PRESSURE: ptr = A.upressure; n = N + S->pcap; break;
    case ALV_FIELD_PID_LENGTH: ptr = A.cell_count; n = S->num_cells; break;
    case ALV_FIELD_GRID_ANOMALY: ptr = A.anomaly; n = 3; break;
    case ALV_FIELD_CELL_START: ptr = A.cell_start; n = size_t(S->num_cells) + 1; break;
    case ALV_FIELD_BOUNDARY_MASK: ptr = A.bmask; n = S->set.num_particles; break;
    case ALV_FIELD_NEIGHBORS:
    case ALV_FIELD_BOUNDARY:
    case ALV_FIELD_BOUNDARY_KERNEL:
    case ALV_FIELD_FORCE:
    case ALV_FIELD_TORQUE:
    case ALV_FIELD_LAST_PRESSURE:
    case ALV_FIELD_AII:
    case ALV_FIELD_ADV_DENSITY:
    case ALV_FIELD_DENSITY_ERR:
    case ALV_FIELD_DII:
    case ALV_FIELD_DIJ_PJ:
    case ALV_FIELD_PRESSURE_ACCEL:
      return export_field(S, field, dptr, n_elements, n_components);
    default: return fail(ALV_ERR_ARG, "unknown field");
  }
  if (need & S->dirty) {
    Launch L;
    int rc = solver_launch(S, &L);
    if (rc != ALV_OK) return rc;
    if (L.active) {
      launch_unsort_fields(L, A, need, S->stream);
      ALV_HIP(hipGetLastError());
    }
    ALV_HIP(hipStreamSynchronize(S->stream));
    S->dirty &= ~need;
  }
  *dptr = ptr;
  if (n_elements) *n_elements = n;
  if (n_components) *n_components = comp;
  return ALV_OK;
}

int alv_solver_get_neighbors(alv_solver* S, unsigned* ids, unsigned* counts) {
  if (!S) return fail(ALV_ERR_ARG, "null solver");
  Launch L;
  int rc = solver_launch(S, &L);
  if (rc != ALV_OK) return rc;
  const size_t K = S->cni.max_num_neighbors_per_particle;
  const size_t n = S->batch ? size_t(S->cap) : size_t(S->set.num_particles);
  if (!S->export_ids) {
    ALV_HIP(hipMalloc(&S->export_ids, K * S->cap * sizeof(unsigned)));
    ALV_HIP(hipMalloc(&S->export_counts, size_t(S->cap) * sizeof(unsigned)));
  }
  ALV_HIP(hipMemsetAsync(S->export_ids, 0xff, K * S->cap * sizeof(unsigned), S->stream));
  ALV_HIP(hipMemsetAsync(S->export_counts, 0, size_t(S->cap) * sizeof(unsigned), S->stream));
  if (L.active) launch_export_neighbors(L, S->A, S->export_ids, S->export_counts, S->stream);
  ALV_HIP(hipGetLastError());
  ALV_HIP(hipStreamSynchronize(S->stream));
  if (ids) ALV_HIP(hipMemcpy(ids, S->export_ids, K * n * sizeof(unsigned), hipMemcpyDeviceToHost));
  if (counts)
    ALV_HIP(hipMemcpy(counts, S->export_counts, n * sizeof(unsigned),
                      hipMemcpyDeviceToHost));
  return ALV_OK;
}

int alv_solver_get_pellet_neighbors(alv_solver* S, int which, unsigned* ids, unsigned* counts) {
  if (!S) return fail(ALV_ERR_ARG, "null solver");
  if (!S->pcap) return fail(ALV_ERR_STATE, "the solver has no pellets");
  if (which != 0 && which != 1) return fail(ALV_ERR_ARG, "which is 0 or 1");
  if (!S->grid_valid)
    return fail(ALV_ERR_STATE, "no neighbour search yet: step or update_particle_neighbors first");
  Launch L;
  int rc = solver_launch(S, &L);
  if (rc != ALV_OK) return rc;
  const DevArrays& A = S->A;
  const unsigned K = S->cni.max_num_neighbors_per_particle, G = nbr_groups(K);
  const unsigned n = S->set.num_particles, np = L.p0.np;
  const unsigned rows = which == 0 ? S->cap : S->pcap, live = which == 0 ? n : np;
  const size_t pspan = size_t(grid_for(S->pcap)) * kBlock;
  std::vector<unsigned> fsorted(S->cap), psorted(pspan), cnt(which == 0 ? S->cap : pspan);
  std::vector<uint4> lists(nbr_capacity(K, rows));
  ALV_HIP(hipStreamSynchronize(S->stream));
  ALV_HIP(hipMemcpy(fsorted.data(), A.sorted_id, fsorted.size() * sizeof(unsigned),
                    hipMemcpyDeviceToHost));
  ALV_HIP(hipMemcpy(psorted.data(), A.psorted_id, psorted.size() * sizeof(unsigned),
                    hipMemcpyDeviceToHost));
  ALV_HIP(hipMemcpy(cnt.data(), which == 0 ? A.pnnb : A.pbnnb, cnt.size() * sizeof(unsigned),
                    hipMemcpyDeviceToHost));
  ALV_HIP(hipMemcpy(lists.data(), which == 0 ? A.pnbr : A.pbnbr, lists.size() * sizeof(uint4),
                    hipMemcpyDeviceToHost));
  if (ids) std::fill(ids, ids + size_t(rows) * K, ~0u);
  if (counts) std::fill(counts, counts + rows, 0u);
  for (unsigned s = 0; s < live; ++s) {
    const unsigned row = which == 0 ? fsorted[s] : psorted[s];
    if (row >= rows) return fail(ALV_ERR_STATE, "pellet lists are out of date");
    const unsigned c = cnt[s];
    if (counts) counts[row] = c;
    if (!ids) continue;
    for (unsigned t = 0; t < c && t < K; ++t) {
      const uint4& g = lists[(size_t(s >> 6) * G + (t >> 2)) * 64u + (s & 63u)];
      const unsigned j = (&g.x)[t & 3u];
      const std::vector<unsigned>& map = which == 0 ? psorted : fsorted;
      if (j >= map.size()) return fail(ALV_ERR_STATE, "pellet lists are out of date");
      ids[size_t(row) * K + t] = which == 0 ? S->cap + map[j] : map[j];
    }
  }
  return ALV_OK;
}

int alv_solver_get_grid(alv_solver* S, unsigned* cell_count, unsigned* cell_start,
                        unsigned* anomaly) {
  if (!S) return fail(ALV_ERR_ARG, "null solver");
  ALV_HIP(hipStreamSynchronize(S->stream));
  const size_t C = S->num_cells;
  if (cell_count)
    ALV_HIP(hipMemcpy(cell_count, S->A.cell_count, C * sizeof(unsigned), hipMemcpyDeviceToHost));
  if (cell_start)
    ALV_HIP(hipMemcpy(cell_start, S->A.cell_start, (C + 1) * sizeof(unsigned),
                      hipMemcpyDeviceToHost));
  if (anomaly) ALV_HIP(hipMemcpy(anomaly, S->A.anomaly, 3 * sizeof(unsigned), hipMemcpyDeviceToHost));
  return ALV_OK;
}

int alv_node_positions(const float dmin[3], const unsigned res[3], const float cs[3],
                       float* out, unsigned* num_nodes) {
  if (!dmin || !res || !cs) return fail(ALV_ERR_ARG, "null argument");
  GridGeomDev g;
  memset(&g, 0, sizeof(g));
  for (int a = 0; a < 3; ++a) {
    g.dmin[a] = dmin[a];
    g.cs[a] = cs[a];
    g.res[a] = res[a];
  }
  size_t nv = size_t(res[0] + 1) * (res[1] + 1) * (res[2] + 1);
  size_t ne = size_t(res[0]) * (res[1] + 1) * (res[2] + 1) +
              size_t(res[0] + 1) * res[1] * (res[2] + 1) +
              size_t(res[0] + 1) * (res[1] + 1) * res[2];
  g.num_nodes = unsigned(nv + 2 * ne);
  if (num_nodes) *num_nodes = g.num_nodes;
  if (!out) return ALV_OK;
  float* d = nullptr;
  ALV_HIP(hipMalloc(&d, size_t(g.num_nodes) * 3 * sizeof(float)));
  launch_node_positions(g, d, 0);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpy(out, d, size_t(g.num_nodes) * 3 * sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(ALV_ERR_HIP, hipGetErrorString(e));
  return ALV_OK;
}

int alv_check_math(const alv_const* cn, int op, const float* a, const float* b, float* out,
                   size_t n) {
  if (!cn || op < 0 || op > 6) return fail(ALV_ERR_ARG, "bad argument");
  if (n && (!a || !out || ((op == 1 || op == 5) && !b)))
    return fail(ALV_ERR_ARG, "null argument");
  if (!n) return ALV_OK;
  alv_consti cni;
  memset(&cni, 0, sizeof(cni));
  SimParams p = make_params(*cn, cni, 0, 0);
  float* d = nullptr;
  ALV_HIP(hipMalloc(&d, 3 * n * sizeof(float)));
  hipError_t e = hipMemcpy(d, a, n * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess && b) e = hipMemcpy(d + n, b, n * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    launch_check_math(p, op, d, d + n, d + 2 * n, n, 0);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(out, d + 2 * n, n * sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(ALV_ERR_HIP, hipGetErrorString(e));
  return ALV_OK;
}

int alv_volume_from_distance_map(const alv_const* cn, const float dmin[3], const float dmax[3],
                                 const unsigned res[3], const float cs[3], float sign,
                                 const float* distance, float* volume, unsigned num_nodes) {
  if (!cn || !distance || !volume) return fail(ALV_ERR_ARG, "null argument");
  GridGeomDev g;
  memset(&g, 0, sizeof(g));
  for (int a = 0; a < 3; ++a) {
    g.dmin[a] = dmin[a];
    g.dmax[a] = dmax[a];
    g.cs[a] = cs[a];
    g.res[a] = res[a];
  }
  g.num_nodes = num_nodes;
  float table[120];
  peirce_table(table);
  float *dd = nullptr, *dv = nullptr, *dp = nullptr;
  ALV_HIP(hipMalloc(&dd, size_t(num_nodes) * sizeof(float)));
  ALV_HIP(hipMalloc(&dv, size_t(num_nodes) * sizeof(float)));
  ALV_HIP(hipMalloc(&dp, sizeof(table)));
  hipError_t e = hipMemcpy(dd, distance, size_t(num_nodes) * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dp, table, sizeof(table), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    alv_consti cni;
    memset(&cni, 0, sizeof(cni));
    SimParams p = make_params(*cn, cni, 0, 0);
    launch_volume_field_grid(p, g, dd, dv, sign, dp, 0, num_nodes, 0);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpy(volume, dv, size_t(num_nodes) * sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(dd);
  (void)hipFree(dv);
  (void)hipFree(dp);
  if (e != hipSuccess) return fail(ALV_ERR_HIP, hipGetErrorString(e));
  return ALV_OK;
}

int alv_profiling_enable(int enable) {
  if (enable && !g_event_profiler) {
    g_event_profiler = new EventProfiler();
    g_profiler = g_event_profiler;
  } else if (!enable && g_event_profiler) {
    ALV_HIP(hipDeviceSynchronize());
    g_profiler = nullptr;
    g_event_profiler->collect();
  }
  return ALV_OK;
}

int alv_profiling_reset(void) {
  if (g_event_profiler) {
    ALV_HIP(hipDeviceSynchronize());
    g_event_profiler->reset();
  }
  return ALV_OK;
}

int alv_profiling_num_kernels(unsigned* n) {
  if (!n) return fail(ALV_ERR_ARG, "null argument");
  if (g_event_profiler) {
    ALV_HIP(hipDeviceSynchronize());
    g_event_profiler->collect();
  }
  *n = g_event_profiler ? unsigned(g_event_profiler->stats().size()) : 0u;
  return ALV_OK;
}

int alv_profiling_kernel(unsigned index, const char** name, double* total_ms,
                         unsigned long long* launches) {
  if (!g_event_profiler || index >= g_event_profiler->stats().size())
    return fail(ALV_ERR_ARG, "kernel index out of range");
  const auto& st = g_event_profiler->stats()[index];
  if (name) *name = st.name.c_str();
  if (total_ms) *total_ms = st.total_ms;
  if (launches) *launches = st.launches;
  return ALV_OK;
}

// ---- sampling and statistics --------------------------------------------

int alv_solver_sample_neighbors(alv_solver* S, const float* sample_x, unsigned ns,
                                unsigned max_neighbors, float* sample_neighbors,
                                unsigned* counts) {
  if (!S || (ns && (!sample_x || !sample_neighbors || !counts)))
    return fail(ALV_ERR_ARG, "null argument");
  if (max_neighbors != S->cni.max_num_neighbors_per_particle)
    return fail(ALV_ERR_ARG, "sample lists must have max_num_neighbors_per_particle entries per sample");
  if (S->batch) return fail(ALV_ERR_UNSUPPORTED, "sampling a batch solver is not supported");
  if (!S->grid_valid)
    return fail(ALV_ERR_STATE, "no neighbour grid yet: step or update_particle_neighbors first");
  ALV_HIP(hipStreamSynchronize(S->stream));
  SimParams p = make_params(S->cn, S->cni, unsigned(S->pile->bnd.size()), S->set.num_particles);
  p.pellet_mode = S->par_host[0].pellet_mode;  // the grid of the last search has pellets
  launch_sample_neighbors(p, S->A, sample_x, ns, reinterpret_cast<float4*>(sample_neighbors),
                          counts, nullptr);
  ALV_HIP(hipGetLastError());
  return ALV_OK;
}

int alv_solver_sample_bead_pellet_neighbors(alv_solver* S, const float* sample_x,
                                            unsigned num_samples, unsigned offset,
                                            unsigned max_num_beads, unsigned max_neighbors,
                                            float* bead_neighbors, unsigned* bead_counts,
                                            float* pellet_neighbors, unsigned* pellet_counts) {
  if (!S || (num_samples && (!sample_x || !bead_neighbors || !bead_counts || !pellet_neighbors ||
                             !pellet_counts)))
    return fail(ALV_ERR_ARG, "null argument");
  if (max_neighbors != S->cni.max_num_neighbors_per_particle)
    return fail(ALV_ERR_ARG, "lists must have max_num_neighbors_per_particle entries per row");
  if (S->batch) return fail(ALV_ERR_UNSUPPORTED, "sampling a batch solver is not supported");
  if (!S->grid_valid)
    return fail(ALV_ERR_STATE, "no neighbour grid yet: step or update_particle_neighbors first");
  ALV_HIP(hipStreamSynchronize(S->stream));
  SimParams p = make_params(S->cn, S->cni, unsigned(S->pile->bnd.size()), S->set.num_particles);
  p.pellet_mode = S->par_host[0].pellet_mode;
  launch_sample_bead_pellet(p, S->A, sample_x, num_samples, offset, max_num_beads,
                            reinterpret_cast<float4*>(bead_neighbors), bead_counts,
                            reinterpret_cast<float4*>(pellet_neighbors), pellet_counts, nullptr);
  ALV_HIP(hipGetLastError());
  return ALV_OK;
}

namespace {
int pellet_sums(const alv_const* cn, int what, const float* nbr, const unsigned* counts,
                const float* pnbr, const unsigned* pcounts, unsigned K, unsigned n, float* out) {
  if (!cn || (n && (!pnbr || !pcounts || !out || (what != 0 && (!nbr || !counts)))))
    return fail(ALV_ERR_ARG, "null argument");
  if (K == 0) return fail(ALV_ERR_ARG, "max_neighbors must be positive");
  alv_consti cni;
  memset(&cni, 0, sizeof(cni));
  SimParams p = make_params(*cn, cni, 0, 0);
  launch_pellet_sums(p, what, reinterpret_cast<const float4*>(nbr), counts,
                     reinterpret_cast<const float4*>(pnbr), pcounts, K, n, out, nullptr);
  ALV_HIP(hipGetLastError());
  return ALV_OK;
}
}  // namespace

int alv_sample_with_pellets(const alv_const* cn, int what, const float* neighbors,
                            const unsigned* counts, const float* pellet_neighbors,
                            const unsigned* pellet_counts, unsigned K, unsigned num_samples,
                            const float* particle_density, const float* particle_v,
                            const float* sample_v, float* out) {
  if (!cn || what < 0 || what > 2) return fail(ALV_ERR_ARG, "bad argument");
  if (num_samples && (!neighbors || !counts || !pellet_neighbors || !pellet_counts || !out ||
                      (what > 0 && (!particle_density || !particle_v)) ||
                      (what == 2 && !sample_v)))
    return fail(ALV_ERR_ARG, "null argument");
  if (K == 0) return fail(ALV_ERR_ARG, "max_neighbors must be positive");
  alv_consti cni;
  memset(&cni, 0, sizeof(cni));
  SimParams p = make_params(*cn, cni, 0, 0);
  launch_sample_with_pellets(p, what, reinterpret_cast<const float4*>(neighbors), counts,
                             reinterpret_cast<const float4*>(pellet_neighbors), pellet_counts, K,
                             num_samples, particle_density, particle_v, sample_v, out, nullptr);
  ALV_HIP(hipGetLastError());
  return ALV_OK;
}

int alv_particle_boundary_with_pellets(const alv_const* cn, const float* pellet_neighbors,
                                       const unsigned* pellet_counts, unsigned K, unsigned n,
                                       float* boundary_kernel) {
  return pellet_sums(cn, 0, nullptr, nullptr, pellet_neighbors, pellet_counts, K, n,
                     boundary_kernel);
}

int alv_density_with_pellets(const alv_const* cn, const float* neighbors, const unsigned* counts,
                             const float* pellet_neighbors, const unsigned* pellet_counts,
                             unsigned K, unsigned n, float* density) {
  return pellet_sums(cn, 1, neighbors, counts, pellet_neighbors, pellet_counts, K, n, density);
}

int alv_dfsph_factor_with_pellets(const alv_const* cn, const float* neighbors,
                                  const unsigned* counts, const float* pellet_neighbors,
                                  const unsigned* pellet_counts, unsigned K, unsigned n,
                                  float* factor) {
  return pellet_sums(cn, 2, neighbors, counts, pellet_neighbors, pellet_counts, K, n, factor);
}

int alv_solver_sample_all_boundaries(alv_solver* S, const float* sample_x, unsigned ns,
                                     float* sample_boundary, float* sample_boundary_kernel) {
  if (!S || (ns && (!sample_x || !sample_boundary || !sample_boundary_kernel)))
    return fail(ALV_ERR_ARG, "null argument");
  if (S->batch) return fail(ALV_ERR_UNSUPPORTED, "sampling a batch solver is not supported");
  alv_pile* P = S->pile;
  const unsigned nb = unsigned(P->bnd.size());
  if (nb == 0 || ns == 0) return ALV_OK;
  ALV_HIP(reserve_dev(&P->bodies_dev, &P->bodies_cap, nb));
  ALV_HIP(hipMemcpy(P->bodies_dev, P->bnd.data(), nb * sizeof(BoundaryDev), hipMemcpyHostToDevice));
  SimParams p = make_params(S->cn, S->cni, nb, S->set.num_particles);
  launch_sample_boundary(p, P->bodies_dev, nb, sample_x, ns,
                         reinterpret_cast<float4*>(sample_boundary),
                         reinterpret_cast<float4*>(sample_boundary_kernel), nullptr);
  ALV_HIP(hipGetLastError());
  ALV_HIP(hipDeviceSynchronize());  // the body table is reused by later calls
  return ALV_OK;
}

int alv_sample_fluid(const alv_const* cn, const float* nbr, const unsigned* counts, unsigned K,
                     unsigned ns, const float* density, const float* q, unsigned comps,
                     float* out) {
  if (!cn || (ns && (!nbr || !counts || !density || !q || !out)))
    return fail(ALV_ERR_ARG, "null argument");
  if (comps != 1 && comps != 3) return fail(ALV_ERR_ARG, "components must be 1 or 3");
  if (K == 0) return fail(ALV_ERR_ARG, "max_neighbors must be positive");
  alv_consti cni;
  memset(&cni, 0, sizeof(cni));
  SimParams p = make_params(*cn, cni, 0, 0);
  launch_sample_fluid(p, reinterpret_cast<const float4*>(nbr), counts, K, ns, density, q, comps,
                      out, nullptr);
  ALV_HIP(hipGetLastError());
  return ALV_OK;
}

int alv_sample_fluid_density(const alv_const* cn, const float* nbr, const unsigned* counts,
                             unsigned K, unsigned ns, float* out) {
  if (!cn || (ns && (!nbr || !counts || !out))) return fail(ALV_ERR_ARG, "null argument");
  if (K == 0) return fail(ALV_ERR_ARG, "max_neighbors must be positive");
  alv_consti cni;
  memset(&cni, 0, sizeof(cni));
  SimParams p = make_params(*cn, cni, 0, 0);
  launch_sample_density(p, reinterpret_cast<const float4*>(nbr), counts, K, ns, nullptr, 0, out,
                        nullptr);
  ALV_HIP(hipGetLastError());
  return ALV_OK;
}

int alv_sample_position_density(const alv_const* cn, const float* nbr, const unsigned* counts,
                                unsigned K, unsigned ns, const float* ker, unsigned nb,
                                float* out) {
  if (!cn || (ns && (!nbr || !counts || !out)) || (nb && ns && !ker))
    return fail(ALV_ERR_ARG, "null argument");
  if (K == 0) return fail(ALV_ERR_ARG, "max_neighbors must be positive");
  alv_consti cni;
  memset(&cni, 0, sizeof(cni));
  SimParams p = make_params(*cn, cni, nb, 0);
  launch_sample_density(p, reinterpret_cast<const float4*>(nbr), counts, K, ns,
                        reinterpret_cast<const float4*>(ker), nb, out, nullptr);
  ALV_HIP(hipGetLastError());
  return ALV_OK;
}

namespace {
int sample_velocity_common(const alv_const* cn, const float* nbr, const unsigned* counts,
                           unsigned K, unsigned ns, const float* density, const float* v,
                           const float* sample_v, const float* rec, const float* ker,
                           unsigned nb, const float* rx, const float* rv, const float* rw,
                           float* out) {
  if (!cn || (ns && (!nbr || !counts || !density || !v || !out)))
    return fail(ALV_ERR_ARG, "null argument");
  if (nb && (!rec || !ker || !rx || !rv || !rw)) return fail(ALV_ERR_ARG, "null boundary argument");
  if (K == 0) return fail(ALV_ERR_ARG, "max_neighbors must be positive");
  alv_consti cni;
  memset(&cni, 0, sizeof(cni));
  SimParams p = make_params(*cn, cni, nb, 0);
  launch_sample_velocity(p, reinterpret_cast<const float4*>(nbr), counts, K, ns, density, v,
                         sample_v, reinterpret_cast<const float4*>(rec),
                         reinterpret_cast<const float4*>(ker), nb, rx, rv, rw, out, nullptr);
  ALV_HIP(hipGetLastError());
  return ALV_OK;
}
}  // namespace

int alv_sample_velocity(const alv_const* cn, const float* nbr, const unsigned* counts, unsigned K,
                        unsigned ns, const float* density, const float* v, const float* rec,
                        const float* ker, unsigned nb, const float* rx, const float* rv,
                        const float* rw, float* out) {
  return sample_velocity_common(cn, nbr, counts, K, ns, density, v, nullptr, rec, ker, nb, rx, rv,
                                rw, out);
}

int alv_sample_vorticity(const alv_const* cn, const float* nbr, const unsigned* counts,
                         unsigned K, unsigned ns, const float* density, const float* v,
                         const float* sample_v, const float* rec, const float* ker, unsigned nb,
                         const float* rx, const float* rv, const float* rw, float* out) {
  if (ns && !sample_v) return fail(ALV_ERR_ARG, "null sample_v");
  return sample_velocity_common(cn, nbr, counts, K, ns, density, v, sample_v, rec, ker, nb, rx, rv,
                                rw, out);
}

int alv_histogram256(const float* v, const float* mask, unsigned n, float v_min, float v_max,
                     unsigned* histogram) {
  if (!histogram || (n && !v)) return fail(ALV_ERR_ARG, "null argument");
  static unsigned* partial = nullptr;  // process-wide scratch of the partial histograms
  if (!partial) ALV_HIP(hipMalloc(&partial, histogram_partial_count() * sizeof(unsigned)));
  const float step_inverse = float(256.0 / (double(v_max) - double(v_min)));
  launch_histogram256(v, mask, n, v_min, step_inverse, partial, histogram, nullptr);
  ALV_HIP(hipGetLastError());
  return ALV_OK;
}

int alv_kl_divergence(const unsigned* hp, const unsigned* hq, unsigned n_p, unsigned n_q,
                      float q_lower_bound, unsigned bins, float* result) {
  if (!hp || !hq || !result) return fail(ALV_ERR_ARG, "null argument");
  if (n_p == 0) return fail(ALV_ERR_ARG, "n_p must be positive");
  float* d = nullptr;
  ALV_HIP(hipMalloc(&d, sizeof(float)));
  launch_kl_divergence(hp, hq, n_p, n_q, q_lower_bound, bins, d, nullptr);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpy(result, d, sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(ALV_ERR_HIP, hipGetErrorString(e));
  return ALV_OK;
}

}  // extern "C"
