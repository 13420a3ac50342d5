// alluvion.hpp -- header-only C++ mirror of the reference's host API for the
// solver path (kennychufk/alluvion: store.hpp, variable.hpp, constants.hpp,
// pile.hpp, solver.hpp, solver_df.hpp, solver_ii.hpp, dg/*_distance.hpp),
// implemented over the C ABI in alv.h.  Class, method and field names follow
// the reference so that host code written against it compiles with few
// changes.  Only TF = float is provided: the device path computes in FP32.
// Errors throw std::runtime_error like the reference (allocator.cu:35-41).
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <functional>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "alv.h"

namespace alluvion {

using U = unsigned int;
using I = int;

struct float2 {
  float x, y;
};
struct float3 {
  float x, y, z;
};
struct float4 {
  float x, y, z, w;
};
struct uint3 {
  U x, y, z;
};
struct int3 {
  I x, y, z;
};
using U3 = uint3;
using I3 = int3;

inline void check(int rc) {
  if (rc != ALV_OK) throw std::runtime_error(alv_last_error());
}

// NumericType (data_type.hpp)
enum class NumericType { f32, f64, i32, u32, undefined };
template <typename M>
struct TypeInfo;
template <>
struct TypeInfo<float> {
  static constexpr NumericType type = NumericType::f32;
  static constexpr U n = 1;
};
template <>
struct TypeInfo<float2> {
  static constexpr NumericType type = NumericType::f32;
  static constexpr U n = 2;
};
template <>
struct TypeInfo<float3> {
  static constexpr NumericType type = NumericType::f32;
  static constexpr U n = 3;
};
template <>
struct TypeInfo<float4> {
  static constexpr NumericType type = NumericType::f32;
  static constexpr U n = 4;
};
template <>
struct TypeInfo<U> {
  static constexpr NumericType type = NumericType::u32;
  static constexpr U n = 1;
};
template <>
struct TypeInfo<I> {
  static constexpr NumericType type = NumericType::i32;
  static constexpr U n = 1;
};
template <>
struct TypeInfo<uint3> {
  static constexpr NumericType type = NumericType::u32;
  static constexpr U n = 3;
};

// Variable<D, M> (variable.hpp:30-316): non-owning {shape, device pointer}.
template <U D, typename M>
class Variable {
 public:
  Variable() : ptr_(nullptr) { shape_.fill(0); }
  Variable(std::array<U, D> const& shape, void* ptr) : shape_(shape), ptr_(ptr) {}
  U get_linear_shape() const {
    U n = 1;
    for (U s : shape_) n *= s;
    return n;
  }
  U get_num_primitives_per_element() const { return TypeInfo<M>::n; }
  U get_num_primitives() const { return get_linear_shape() * TypeInfo<M>::n; }
  U get_num_bytes() const { return get_linear_shape() * sizeof(M); }
  NumericType get_type() const { return TypeInfo<M>::type; }
  std::vector<U> get_shape() const { return std::vector<U>(shape_.begin(), shape_.end()); }
  // Offsets count elements of M, as in the reference (variable.hpp:60-82).
  void get_bytes(void* dst, U num_bytes, U offset = 0) const {
    if (num_bytes == 0) return;
    const size_t byte_offset = size_t(offset) * sizeof(M);
    if (byte_offset + num_bytes > get_num_bytes())
      throw std::out_of_range("Retrieving more bytes than allocated");
    check(alv_memcpy_d2h(dst, static_cast<char*>(ptr_) + byte_offset, num_bytes));
  }
  void get_bytes(void* dst) const { get_bytes(dst, get_num_bytes()); }
  void set_bytes(void const* src, U num_bytes, U offset = 0) {
    if (num_bytes == 0) return;
    const size_t byte_offset = size_t(offset) * sizeof(M);
    if (byte_offset + num_bytes > get_num_bytes())
      throw std::out_of_range("Setting more bytes than allocated");
    check(alv_memcpy_h2d(static_cast<char*>(ptr_) + byte_offset, src, num_bytes));
  }
  void set_bytes(void const* src) { set_bytes(src, get_num_bytes()); }
  void set_zero() { check(alv_memset(ptr_, 0, get_num_bytes())); }
  void set_same(int value, U num_elements, U offset = 0) {
    if (num_elements == 0) return;
    if ((size_t(offset) + num_elements) * sizeof(M) > get_num_bytes())
      throw std::out_of_range("Setting more bytes than allocated");
    check(alv_memset(static_cast<char*>(ptr_) + offset * sizeof(M), value,
                     num_elements * sizeof(M)));
  }
  void set_same(int value) { check(alv_memset(ptr_, value, get_num_bytes())); }
  void set_from(Variable const& src, U num_elements, U offset = 0) {
    if (num_elements == 0) return;
    if ((size_t(offset) + num_elements) * sizeof(M) > get_num_bytes() ||
        size_t(num_elements) * sizeof(M) > src.get_num_bytes())
      throw std::out_of_range("Copying more bytes than allocated");
    check(alv_memcpy_d2d(static_cast<char*>(ptr_) + offset * sizeof(M), src.ptr_,
                         num_elements * sizeof(M)));
  }
  void set_from(Variable const& src) {
    check(alv_memcpy_d2d(ptr_, src.ptr_, std::min(get_num_bytes(), src.get_num_bytes())));
  }

  // .alu files (variable.hpp:132-261): D u32 shape items (the outermost
  // may be smaller than the variable's), a u32 0, the u32 number of
  // primitives per element, a one-byte type label ('f', 'd', 'i', 'u')
  // and the raw little-endian elements.  write_file writes
  // `shape_outermost` outer rows starting at element `offset`; 0 means
  // the whole outermost extent (the reference sizes the payload before
  // applying that default and so writes no data then; here the payload
  // always matches the header).
  void write_file(const char* filename, U shape_outermost = 0, U offset = 0) const {
    if (shape_outermost > shape_[0])
      throw std::runtime_error("Writing more than the outermost shape (" +
                               std::to_string(shape_outermost) + ">" +
                               std::to_string(shape_[0]) + ").");
    if (shape_outermost == 0) shape_outermost = shape_[0];
    const size_t inner = shape_[0] ? get_linear_shape() / shape_[0] : 0;
    const size_t linear = inner * shape_outermost;
    const size_t num_bytes = linear * sizeof(M);
    if (num_bytes + size_t(offset) * sizeof(M) > get_num_bytes())
      throw std::runtime_error("Writing more than allocated: " +
                               std::to_string(num_bytes + size_t(offset) * sizeof(M)) + " " +
                               std::to_string(get_num_bytes()));
    std::ofstream stream(filename, std::ios::binary | std::ios::trunc);
    if (!stream.is_open()) throw std::runtime_error(std::string("Failed to open alu file: ") + filename);
    for (U i = 0; i < D; ++i) {
      const U item = i == 0 ? shape_outermost : shape_[i];
      stream.write(reinterpret_cast<const char*>(&item), sizeof(U));
    }
    const U end_of_shape = 0, prims = get_num_primitives_per_element();
    stream.write(reinterpret_cast<const char*>(&end_of_shape), sizeof(U));
    stream.write(reinterpret_cast<const char*>(&prims), sizeof(U));
    const char label = type_label(get_type());
    stream.write(&label, 1);
    std::vector<char> host(num_bytes);
    get_bytes(host.data(), U(num_bytes), offset);
    stream.write(host.data(), std::streamsize(num_bytes));
    if (!stream) throw std::runtime_error(std::string("Failed to write alu file: ") + filename);
  }

  // Reads an .alu file into the first rows of the variable and returns the
  // file's outermost extent.  Inner extents, the element width and the type
  // must match; FP32 and FP64 files convert into each other.
  U read_file(const char* filename) {
    std::ifstream stream(filename, std::ios::binary);
    if (!stream.is_open()) throw std::runtime_error(std::string("Failed to open alu file: ") + filename);
    U outermost = 0;
    for (U i = 0; i < D; ++i) {
      U item = 0;
      stream.read(reinterpret_cast<char*>(&item), sizeof(U));
      if (!stream || item == 0)
        throw std::runtime_error(std::string("Shape mismatch when reading ") + filename);
      if (i == 0) {
        outermost = item;
        if (outermost > shape_[0])
          throw std::runtime_error("Reading more than the outermost shape (" +
                                   std::to_string(outermost) + ">" + std::to_string(shape_[0]) +
                                   "): " + filename + ".");
      } else if (item != shape_[i]) {
        throw std::runtime_error("Shape mismatch for shape item " + std::to_string(i) + "(" +
                                 std::to_string(shape_[i]) + "!=" + std::to_string(item) +
                                 "): " + filename + ".");
      }
    }
    U end_of_shape = 1, prims = 0;
    stream.read(reinterpret_cast<char*>(&end_of_shape), sizeof(U));
    if (!stream || end_of_shape != 0)
      throw std::runtime_error(std::string("Dimension mismatch when reading ") + filename);
    stream.read(reinterpret_cast<char*>(&prims), sizeof(U));
    if (!stream || prims != get_num_primitives_per_element())
      throw std::runtime_error(std::string("Num primitives per unit mismatch when reading ") +
                               filename + "(" + std::to_string(prims) +
                               "!=" + std::to_string(get_num_primitives_per_element()) + ").");
    char label = 0;
    stream.read(&label, 1);
    const size_t linear = size_t(get_linear_shape() / shape_[0]) * outermost;
    const size_t num_prims = linear * prims;
    std::vector<M> host(linear);
    const char mine = type_label(get_type());
    if (label == mine) {
      stream.read(reinterpret_cast<char*>(host.data()), std::streamsize(linear * sizeof(M)));
    } else if (label == 'd' && mine == 'f') {
      std::vector<double> src(num_prims);
      stream.read(reinterpret_cast<char*>(src.data()), std::streamsize(num_prims * sizeof(double)));
      float* dst = reinterpret_cast<float*>(host.data());
      for (size_t i = 0; i < num_prims; ++i) dst[i] = static_cast<float>(src[i]);
    } else {
      throw std::runtime_error(std::string("Data type mismatch when reading ") + filename);
    }
    if (!stream) throw std::runtime_error(std::string("Truncated alu file: ") + filename);
    set_bytes(host.data(), U(linear * sizeof(M)));
    return outermost;
  }

  static char type_label(NumericType t) {
    switch (t) {
      case NumericType::f32: return 'f';
      case NumericType::f64: return 'd';
      case NumericType::i32: return 'i';
      case NumericType::u32: return 'u';
      default: return '!';
    }
  }

  std::array<U, D> shape_;
  void* ptr_;
};

// Const<float> (constants.hpp:51-137)
template <typename TF>
struct Const;
template <>
struct Const<float> {
  float kernel_radius = 0, kernel_radius_sqr = 0, cubic_kernel_k = 0, cubic_kernel_l = 0,
        cubic_kernel_zero = 0;
  float particle_radius = 0, particle_vol = 0, particle_mass = 0, density0 = 0;
  float viscosity = 0, boundary_viscosity = 0, vorticity_coeff = 0,
        boundary_vorticity_coeff = 0, inertia_inverse = 0, viscosity_omega = 0,
        surface_tension_coeff = 0, surface_tension_boundary_coeff = 0;
  float3 gravity{0, 0, 0};
  float boundary_epsilon = 0, dfsph_factor_epsilon = 0, contact_tolerance = 0;
  float wrap_length = 0, wrap_min = 0, wrap_max = 0;
  // extension: periodic x and z (0 = off)
  float wrap_length_x = 0, wrap_length_z = 0;

  void set_kernel_radius(float h) {
    alv_const c{};
    alv_const_set_kernel_radius(&c, h);
    kernel_radius = c.kernel_radius;
    kernel_radius_sqr = c.kernel_radius_sqr;
    cubic_kernel_k = c.cubic_kernel_k;
    cubic_kernel_l = c.cubic_kernel_l;
    cubic_kernel_zero = c.cubic_kernel_zero;
  }
  void set_particle_attr(float radius, float mass, float density) {
    alv_const c{};
    alv_const_set_particle_attr(&c, radius, mass, density);
    particle_radius = c.particle_radius;
    particle_mass = c.particle_mass;
    density0 = c.density0;
    particle_vol = c.particle_vol;
  }
  void set_wrap_length(float l) {
    wrap_length = l;
    wrap_min = l * -0.5f;
    wrap_max = l * 0.5f;
  }
  void set_cubic_discretization_constants() {}
  alv_const to_c() const {
    alv_const c{};
    c.kernel_radius = kernel_radius;
    c.kernel_radius_sqr = kernel_radius_sqr;
    c.cubic_kernel_k = cubic_kernel_k;
    c.cubic_kernel_l = cubic_kernel_l;
    c.cubic_kernel_zero = cubic_kernel_zero;
    c.particle_radius = particle_radius;
    c.particle_vol = particle_vol;
    c.particle_mass = particle_mass;
    c.density0 = density0;
    c.viscosity = viscosity;
    c.boundary_viscosity = boundary_viscosity;
    c.gravity[0] = gravity.x;
    c.gravity[1] = gravity.y;
    c.gravity[2] = gravity.z;
    c.boundary_epsilon = boundary_epsilon;
    c.dfsph_factor_epsilon = dfsph_factor_epsilon;
    c.contact_tolerance = contact_tolerance;
    c.wrap_length[0] = wrap_length_x;
    c.wrap_length[1] = wrap_length;
    c.wrap_length[2] = wrap_length_z;
    return c;
  }
};
using Constf = Const<float>;

// Consti (constants.hpp:40-49)
struct ConstiN {
  U max_num_particles_per_cell = 64;
  U3 grid_res{0, 0, 0};
  I3 grid_offset{0, 0, 0};
  U max_num_neighbors_per_particle = 64;
  U num_boundaries = 0;
  U max_num_contacts = 0;
  alv_consti to_c() const {
    alv_consti c{};
    c.max_num_particles_per_cell = max_num_particles_per_cell;
    c.grid_res[0] = grid_res.x;
    c.grid_res[1] = grid_res.y;
    c.grid_res[2] = grid_res.z;
    c.grid_offset[0] = grid_offset.x;
    c.grid_offset[1] = grid_offset.y;
    c.grid_offset[2] = grid_offset.z;
    c.max_num_neighbors_per_particle = max_num_neighbors_per_particle;
    c.num_boundaries = num_boundaries;
    c.max_num_contacts = max_num_contacts;
    return c;
  }
};

// Store (store.hpp:18-170): device buffers, constants.
class Store {
 public:
  Store() = default;
  ~Store() {
    for (void* p : owned_) alv_free(p);
    if (device_constants() == &cn_dev_) device_constants() = nullptr;
  }
  // The constants the device code of the last copy_cn / copy_cn_external
  // uses (the reference keeps them in __constant__ memory,
  // constants.hpp:141-143); Runner's stand-alone launchers read them.
  static Const<float> const*& device_constants() {
    static Const<float> const* current = nullptr;
    return current;
  }
  static void set_device(int device) { check(alv_set_device(device)); }
  template <U D, typename M>
  Variable<D, M>* create(std::array<U, D> const& shape) {
    size_t n = 1;
    for (U s : shape) n *= s;
    void* p = nullptr;
    check(alv_malloc(&p, n * sizeof(M)));
    if (p) owned_.push_back(p);
    return new Variable<D, M>(shape, p);
  }
  template <U D, typename M>
  void remove(Variable<D, M>& var) {
    for (auto it = owned_.begin(); it != owned_.end(); ++it) {
      if (*it == var.ptr_) {
        alv_free(var.ptr_);
        owned_.erase(it);
        break;
      }
    }
    var.ptr_ = nullptr;
  }
  // Store::get_alu_info (store.cu:47-66): an .alu file's shape (outermost
  // first), primitives per element and type label, from its header.
  static std::tuple<std::vector<U>, U, char> get_alu_info(const char* filename) {
    std::tuple<std::vector<U>, U, char> info;
    std::ifstream stream(filename, std::ios::binary);
    if (!stream.is_open()) throw std::runtime_error(std::string("Failed to open alu file: ") + filename);
    for (;;) {
      U item = 0;
      stream.read(reinterpret_cast<char*>(&item), sizeof(U));
      if (!stream) throw std::runtime_error(std::string("Truncated alu header: ") + filename);
      if (item == 0) break;
      std::get<0>(info).push_back(item);
    }
    stream.read(reinterpret_cast<char*>(&std::get<1>(info)), sizeof(U));
    stream.read(&std::get<2>(info), 1);
    if (!stream) throw std::runtime_error(std::string("Truncated alu header: ") + filename);
    return info;
  }
  Const<float>& get_cn() { return cn_; }
  ConstiN& get_cni() { return cni_; }
  std::tuple<Const<float>, ConstiN> create_cn() { return {Const<float>(), ConstiN()}; }
  // copy_cn / copy_cn_external (store.hpp:142-164): the constants that the
  // next solver step will use.
  void copy_cn() {
    cn_dev_ = cn_;
    cni_dev_ = cni_;
    has_dev_ = true;
    device_constants() = &cn_dev_;
  }
  void copy_cn_external(Const<float> const& cn, ConstiN const& cni) {
    cn_dev_ = cn;
    cni_dev_ = cni;
    has_dev_ = true;
    device_constants() = &cn_dev_;
  }
  bool has_device_constants() const { return has_dev_; }
  Const<float> const& device_cn() const { return cn_dev_; }
  ConstiN const& device_cni() const { return cni_dev_; }

 private:
  std::vector<void*> owned_;
  Const<float> cn_, cn_dev_;
  ConstiN cni_, cni_dev_;
  bool has_dev_ = false;
};

// The neighbour grid of a solver (its pid / pid_length pair in the
// reference): what Runner::launch_make_neighbor_list searches.
struct SolverGrid {
  alv_solver* solver = nullptr;
};

// Runner<float> (runner.hpp:4558-6312): reductions.
template <typename TF>
class Runner {
 public:
  template <U D, typename M>
  static TF sum(Variable<D, M> const& var, U n, U offset = 0) {
    double s = 0;
    check(alv_sum_f32(reinterpret_cast<const float*>(static_cast<const char*>(var.ptr_) +
                                                     size_t(offset) * sizeof(M)),
                      size_t(n) * TypeInfo<M>::n, &s));
    return TF(s);
  }
  template <U D, typename M>
  static TF max(Variable<D, M> const& var, U n, U offset = 0) {
    float m = 0;
    check(alv_max_f32(reinterpret_cast<const float*>(static_cast<const char*>(var.ptr_) +
                                                     size_t(offset) * sizeof(M)),
                      size_t(n) * TypeInfo<M>::n, &m));
    return TF(m);
  }

  // ---- sampling and statistics (runner.hpp:4173-4356, 5698-5890) ---------
  // Sample neighbour lists are Variable<2, float4> [num_samples][K].
  // make_neighbor_list for sample points (runner.hpp:5604-5620) on a
  // solver's grid; the periodic wrap is the solver's.
  static void launch_make_neighbor_list(Variable<1, float3> const& sample_x, SolverGrid pid,
                                        Variable<2, float4>& sample_neighbors,
                                        Variable<1, U>& sample_num_neighbors, U num_samples) {
    if (!pid.solver) throw std::runtime_error("pid does not belong to a solver");
    if (sample_x.shape_[0] < num_samples) throw std::out_of_range("sample_x too small");
    check_samples(sample_neighbors, sample_num_neighbors, num_samples);
    check(alv_solver_sample_neighbors(pid.solver, f(sample_x), num_samples,
                                      sample_neighbors.shape_[1],
                                      static_cast<float*>(sample_neighbors.ptr_),
                                      static_cast<unsigned*>(sample_num_neighbors.ptr_)));
  }
  // make_bead_pellet_neighbor_list (runner.hpp:5622-5641)
  static void launch_make_bead_pellet_neighbor_list(
      Variable<1, float3> const& sample_x, SolverGrid pid, Variable<2, float4>& bead_neighbors,
      Variable<1, U>& num_bead_neighbors, Variable<2, float4>& pellet_neighbors,
      Variable<1, U>& num_pellet_neighbors, U max_num_beads, U num_samples, U offset = 0) {
    if (!pid.solver) throw std::runtime_error("pid does not belong to a solver");
    if (sample_x.shape_[0] < offset + num_samples) throw std::out_of_range("sample_x too small");
    check_samples(bead_neighbors, num_bead_neighbors, offset + num_samples);
    check_samples(pellet_neighbors, num_pellet_neighbors, offset + num_samples);
    check(alv_solver_sample_bead_pellet_neighbors(
        pid.solver, f(sample_x), num_samples, offset, max_num_beads, bead_neighbors.shape_[1],
        static_cast<float*>(bead_neighbors.ptr_), static_cast<unsigned*>(num_bead_neighbors.ptr_),
        static_cast<float*>(pellet_neighbors.ptr_),
        static_cast<unsigned*>(num_pellet_neighbors.ptr_)));
  }
  // sample_position_density_with_pellets, compute_sample_velocity_with_
  // pellets, compute_sample_vorticity_with_pellets (runner.hpp:5747-5846)
  static void launch_sample_density_with_pellets(
      Variable<1, float3> const& sample_x, Variable<2, float4> const& nbr,
      Variable<1, U> const& cnt, Variable<2, float4> const& pel, Variable<1, U> const& pcnt,
      Variable<1, float>& sample_density, U num_samples) {
    (void)sample_x;
    pellet_samples(nbr, cnt, pel, pcnt, num_samples);
    if (sample_density.shape_[0] < num_samples) throw std::out_of_range("sample_density too small");
    alv_const c = constants();
    check(alv_sample_with_pellets(&c, 0, f(nbr), u(cnt), f(pel), u(pcnt), nbr.shape_[1],
                                  num_samples, nullptr, nullptr, nullptr,
                                  static_cast<float*>(sample_density.ptr_)));
  }
  static void launch_sample_velocity_with_pellets(
      Variable<1, float3> const& sample_x, Variable<1, float3> const& particle_x,
      Variable<1, float> const& particle_density, Variable<1, float3> const& particle_v,
      Variable<2, float4> const& nbr, Variable<1, U> const& cnt, Variable<1, float3>& sample_v,
      Variable<2, float4> const& pel, Variable<1, U> const& pcnt, U num_samples) {
    (void)sample_x;
    (void)particle_x;
    pellet_samples(nbr, cnt, pel, pcnt, num_samples);
    if (sample_v.shape_[0] < num_samples) throw std::out_of_range("sample_v too small");
    alv_const c = constants();
    check(alv_sample_with_pellets(&c, 1, f(nbr), u(cnt), f(pel), u(pcnt), nbr.shape_[1],
                                  num_samples, f(particle_density), f(particle_v), nullptr,
                                  static_cast<float*>(sample_v.ptr_)));
  }
  static void launch_sample_vorticity_with_pellets(
      Variable<1, float3> const& sample_x, Variable<1, float3> const& particle_x,
      Variable<1, float> const& particle_density, Variable<1, float3> const& particle_v,
      Variable<2, float4> const& nbr, Variable<1, U> const& cnt,
      Variable<1, float3> const& sample_v, Variable<1, float3>& sample_vorticity,
      Variable<2, float4> const& pel, Variable<1, U> const& pcnt, U num_samples) {
    (void)sample_x;
    (void)particle_x;
    pellet_samples(nbr, cnt, pel, pcnt, num_samples);
    if (sample_v.shape_[0] < num_samples || sample_vorticity.shape_[0] < num_samples)
      throw std::out_of_range("sample_v / sample_vorticity too small");
    alv_const c = constants();
    check(alv_sample_with_pellets(&c, 2, f(nbr), u(cnt), f(pel), u(pcnt), nbr.shape_[1],
                                  num_samples, f(particle_density), f(particle_v), f(sample_v),
                                  static_cast<float*>(sample_vorticity.ptr_)));
  }
  static void pellet_samples(Variable<2, float4> const& nbr, Variable<1, U> const& cnt,
                             Variable<2, float4> const& pel, Variable<1, U> const& pcnt, U ns) {
    check_samples(nbr, cnt, ns);
    check_samples(pel, pcnt, ns);
    if (pel.shape_[1] != nbr.shape_[1])
      throw std::invalid_argument("the two lists must have the same row length");
  }
  // compute_particle_boundary_with_pellets (runner.hpp:5693-5703)
  static void launch_compute_particle_boundary_with_pellets(
      Variable<1, float4>& boundary_kernel_combined, Variable<2, float4> const& pellet_neighbors,
      Variable<1, U> const& num_pellet_neighbors, U num_particles) {
    check_samples(pellet_neighbors, num_pellet_neighbors, num_particles);
    if (boundary_kernel_combined.shape_[0] < num_particles)
      throw std::out_of_range("particle_boundary_kernel_combined too small");
    alv_const c = constants();
    check(alv_particle_boundary_with_pellets(&c, f(pellet_neighbors), u(num_pellet_neighbors),
                                             pellet_neighbors.shape_[1], num_particles,
                                             static_cast<float*>(boundary_kernel_combined.ptr_)));
  }
  // compute_density_with_pellets (runner.hpp:5681-5692)
  static void launch_compute_density_with_pellets(
      Variable<2, float4> const& neighbors, Variable<1, U> const& num_neighbors,
      Variable<2, float4> const& pellet_neighbors, Variable<1, U> const& num_pellet_neighbors,
      Variable<1, float>& density, U num_particles) {
    check_samples(neighbors, num_neighbors, num_particles);
    check_samples(pellet_neighbors, num_pellet_neighbors, num_particles);
    if (pellet_neighbors.shape_[1] != neighbors.shape_[1])
      throw std::invalid_argument("the two lists must have the same row length");
    if (density.shape_[0] < num_particles) throw std::out_of_range("particle_density too small");
    alv_const c = constants();
    check(alv_density_with_pellets(&c, f(neighbors), u(num_neighbors), f(pellet_neighbors),
                                   u(num_pellet_neighbors), neighbors.shape_[1], num_particles,
                                   static_cast<float*>(density.ptr_)));
  }
  // compute_dfsph_factor_with_pellets (runner.hpp:3156-3190; launched by
  // SolverPellet, solver_pellet.hpp:77-87)
  static void launch_compute_dfsph_factor_with_pellets(
      Variable<2, float4> const& neighbors, Variable<1, U> const& num_neighbors,
      Variable<2, float4> const& pellet_neighbors, Variable<1, U> const& num_pellet_neighbors,
      Variable<1, float>& factor, U num_particles) {
    check_samples(neighbors, num_neighbors, num_particles);
    check_samples(pellet_neighbors, num_pellet_neighbors, num_particles);
    if (pellet_neighbors.shape_[1] != neighbors.shape_[1])
      throw std::invalid_argument("the two lists must have the same row length");
    if (factor.shape_[0] < num_particles) throw std::out_of_range("particle_dfsph_factor too small");
    alv_const c = constants();
    check(alv_dfsph_factor_with_pellets(&c, f(neighbors), u(num_neighbors), f(pellet_neighbors),
                                        u(num_pellet_neighbors), neighbors.shape_[1],
                                        num_particles, static_cast<float*>(factor.ptr_)));
  }
  template <typename Q>
  static void launch_sample_fluid(Variable<1, float3> const& sample_x,
                                  Variable<1, float3> const& particle_x,
                                  Variable<1, float> const& particle_density,
                                  Variable<1, Q> const& particle_quantity,
                                  Variable<2, float4> const& sample_neighbors,
                                  Variable<1, U> const& sample_num_neighbors,
                                  Variable<1, Q>& sample_quantity, U num_samples) {
    (void)sample_x;
    (void)particle_x;
    check_samples(sample_neighbors, sample_num_neighbors, num_samples);
    if (sample_quantity.shape_[0] < num_samples) throw std::out_of_range("sample_quantity too small");
    alv_const c = constants();
    check(alv_sample_fluid(&c, f(sample_neighbors), u(sample_num_neighbors), sample_neighbors.shape_[1],
                           num_samples, f(particle_density), f(particle_quantity), TypeInfo<Q>::n,
                           static_cast<float*>(sample_quantity.ptr_)));
  }
  static void launch_sample_fluid_density(Variable<2, float4> const& sample_neighbors,
                                          Variable<1, U> const& sample_num_neighbors,
                                          Variable<1, float>& sample_density, U num_samples) {
    check_samples(sample_neighbors, sample_num_neighbors, num_samples);
    if (sample_density.shape_[0] < num_samples) throw std::out_of_range("sample_density too small");
    alv_const c = constants();
    check(alv_sample_fluid_density(&c, f(sample_neighbors), u(sample_num_neighbors),
                                   sample_neighbors.shape_[1], num_samples,
                                   static_cast<float*>(sample_density.ptr_)));
  }
  static void launch_sample_density(Variable<1, float3> const& sample_x,
                                    Variable<2, float4> const& sample_neighbors,
                                    Variable<1, U> const& sample_num_neighbors,
                                    Variable<1, float>& sample_density,
                                    Variable<2, float4> const& sample_boundary_kernel,
                                    U num_samples) {
    (void)sample_x;
    check_samples(sample_neighbors, sample_num_neighbors, num_samples);
    const U nb = sample_boundary_kernel.shape_[0];
    if (nb && sample_boundary_kernel.shape_[1] != num_samples)
      throw std::out_of_range("sample_boundary_kernel must be [num_boundaries][num_samples]");
    if (sample_density.shape_[0] < num_samples) throw std::out_of_range("sample_density too small");
    alv_const c = constants();
    check(alv_sample_position_density(&c, f(sample_neighbors), u(sample_num_neighbors),
                                      sample_neighbors.shape_[1], num_samples,
                                      f(sample_boundary_kernel), nb,
                                      static_cast<float*>(sample_density.ptr_)));
  }
  static void launch_sample_velocity(
      Variable<1, float3> const& sample_x, Variable<1, float3> const& particle_x,
      Variable<1, float> const& particle_density, Variable<1, float3> const& particle_v,
      Variable<2, float4> const& sample_neighbors, Variable<1, U> const& sample_num_neighbors,
      Variable<1, float3>& sample_v, Variable<2, float4> const& sample_boundary,
      Variable<2, float4> const& sample_boundary_kernel, Variable<1, float3> const& rigid_x,
      Variable<1, float3> const& rigid_v, Variable<1, float3> const& rigid_omega, U num_samples) {
    (void)sample_x;
    (void)particle_x;
    check_samples(sample_neighbors, sample_num_neighbors, num_samples);
    const U nb = boundaries(sample_boundary, sample_boundary_kernel, rigid_x, rigid_v, rigid_omega,
                            num_samples);
    alv_const c = constants();
    check(alv_sample_velocity(&c, f(sample_neighbors), u(sample_num_neighbors),
                              sample_neighbors.shape_[1], num_samples, f(particle_density),
                              f(particle_v), f(sample_boundary), f(sample_boundary_kernel), nb,
                              f(rigid_x), f(rigid_v), f(rigid_omega),
                              static_cast<float*>(sample_v.ptr_)));
  }
  static void launch_sample_vorticity(
      Variable<1, float3> const& sample_x, Variable<1, float3> const& particle_x,
      Variable<1, float> const& particle_density, Variable<1, float3> const& particle_v,
      Variable<2, float4> const& sample_neighbors, Variable<1, U> const& sample_num_neighbors,
      Variable<1, float3> const& sample_v, Variable<1, float3>& sample_vorticity,
      Variable<2, float4> const& sample_boundary, Variable<2, float4> const& sample_boundary_kernel,
      Variable<1, float3> const& rigid_x, Variable<1, float3> const& rigid_v,
      Variable<1, float3> const& rigid_omega, U num_samples) {
    (void)sample_x;
    (void)particle_x;
    check_samples(sample_neighbors, sample_num_neighbors, num_samples);
    const U nb = boundaries(sample_boundary, sample_boundary_kernel, rigid_x, rigid_v, rigid_omega,
                            num_samples);
    alv_const c = constants();
    check(alv_sample_vorticity(&c, f(sample_neighbors), u(sample_num_neighbors),
                               sample_neighbors.shape_[1], num_samples, f(particle_density),
                               f(particle_v), f(sample_v), f(sample_boundary),
                               f(sample_boundary_kernel), nb, f(rigid_x), f(rigid_v),
                               f(rigid_omega), static_cast<float*>(sample_vorticity.ptr_)));
  }
  // The partial histogram and quantized buffers of the reference's signature
  // are not needed here.
  static void launch_histogram256(Variable<1, float> const& v, float v_min, float v_max, U n,
                                  Variable<1, U>& histogram, Variable<1, float> const* mask = nullptr) {
    if (v.get_linear_shape() < n || histogram.get_linear_shape() < 256 ||
        (mask && mask->get_linear_shape() < n))
      throw std::out_of_range("histogram256: buffers too small");
    check(alv_histogram256(f(v), mask ? f(*mask) : nullptr, n, v_min, v_max,
                           static_cast<unsigned*>(histogram.ptr_)));
  }
  static TF calculate_kl_divergence(Variable<1, U> const& histogram_p,
                                    Variable<1, U> const& histogram_q, U n_p, U n_q,
                                    TF q_lower_bound, U num_bins) {
    if (histogram_p.get_linear_shape() < num_bins || histogram_q.get_linear_shape() < num_bins)
      throw std::out_of_range("histograms hold fewer than num_bins bins");
    float r = 0;
    check(alv_kl_divergence(u(histogram_p), u(histogram_q), n_p, n_q, q_lower_bound, num_bins, &r));
    return TF(r);
  }

 private:
  template <U D, typename M>
  static const float* f(Variable<D, M> const& v) {
    return static_cast<const float*>(v.ptr_);
  }
  template <U D, typename M>
  static const unsigned* u(Variable<D, M> const& v) {
    return static_cast<const unsigned*>(v.ptr_);
  }
  static alv_const constants() {
    Const<float> const* c = Store::device_constants();
    if (!c) throw std::runtime_error("no device constants: create a solver or call copy_cn first");
    return c->to_c();
  }
  static void check_samples(Variable<2, float4> const& nbr, Variable<1, U> const& counts, U ns) {
    if (nbr.shape_[0] < ns || counts.shape_[0] < ns)
      throw std::out_of_range("sample buffers hold fewer than num_samples samples");
  }
  static U boundaries(Variable<2, float4> const& rec, Variable<2, float4> const& ker,
                      Variable<1, float3> const& rx, Variable<1, float3> const& rv,
                      Variable<1, float3> const& rw, U ns) {
    const U nb = rec.shape_[0];
    if (nb && (rec.shape_[1] != ns || ker.shape_[0] != nb || ker.shape_[1] != ns ||
               rx.shape_[0] < nb || rv.shape_[0] < nb || rw.shape_[0] < nb))
      throw std::out_of_range("sample boundary buffers must be [num_boundaries][num_samples]");
    return nb;
  }
};

namespace dg {
// Distance hierarchy (dg/distance.hpp, box_distance.hpp,
// box_shell_distance.hpp, sphere_distance.hpp).
template <typename TF3, typename TF>
class Distance {
 public:
  virtual ~Distance() = default;
  virtual alv_distance to_c() const = 0;
  TF3 aabb_min_{0, 0, 0}, aabb_max_{0, 0, 0};
  TF max_distance_ = 0;
  TF3 get_aabb_min() const { return aabb_min_; }
  TF3 get_aabb_max() const { return aabb_max_; }
  TF get_max_distance() const { return max_distance_; }
};
template <typename TF3, typename TF>
class BoxDistance : public Distance<TF3, TF> {
 public:
  BoxDistance(TF3 widths, TF outset = 0) : widths_(widths), outset_(outset) {
    this->aabb_min_ = {widths.x * -0.5f - outset, widths.y * -0.5f - outset,
                       widths.z * -0.5f - outset};
    this->aabb_max_ = {widths.x * 0.5f + outset, widths.y * 0.5f + outset,
                       widths.z * 0.5f + outset};
  }
  static BoxDistance* create(TF3 widths, TF outset = 0) { return new BoxDistance(widths, outset); }
  alv_distance to_c() const override {
    alv_distance d{};
    d.type = ALV_DISTANCE_BOX;
    d.params[0] = widths_.x;
    d.params[1] = widths_.y;
    d.params[2] = widths_.z;
    d.params[3] = outset_;
    return d;
  }
  TF3 widths_;
  TF outset_;
};
template <typename TF3, typename TF>
class BoxShellDistance : public Distance<TF3, TF> {
 public:
  BoxShellDistance(TF3 inner_widths, TF thickness, TF outset = 0)
      : inner_(inner_widths), thickness_(thickness), outset_(outset) {}
  static BoxShellDistance* create(TF3 inner_widths, TF thickness, TF outset = 0) {
    return new BoxShellDistance(inner_widths, thickness, outset);
  }
  alv_distance to_c() const override {
    alv_distance d{};
    d.type = ALV_DISTANCE_BOX_SHELL;
    d.params[0] = inner_.x;
    d.params[1] = inner_.y;
    d.params[2] = inner_.z;
    d.params[3] = thickness_;
    d.params[4] = outset_;
    return d;
  }
  TF3 inner_;
  TF thickness_, outset_;
};
template <typename TF3, typename TF>
class SphereDistance : public Distance<TF3, TF> {
 public:
  explicit SphereDistance(TF radius) : radius_(radius) {}
  static SphereDistance* create(TF radius) { return new SphereDistance(radius); }
  alv_distance to_c() const override {
    alv_distance d{};
    d.type = ALV_DISTANCE_SPHERE;
    d.params[0] = radius_;
    return d;
  }
  TF radius_;
};
// A distance sampled on the boundary's map nodes, as the reference does for
// MeshDistance (construct_distance_grid, pile.hpp:79-95): `fn` writes the
// signed distance of n points (xyz triples, body frame) to out[n]; the
// boundary records, the volume map and the contacts then use the
// interpolated map and its gradient.
template <typename TF3, typename TF>
class MapDistance : public Distance<TF3, TF> {
 public:
  using Fn = std::function<void(const float* points, unsigned n, float* out)>;
  MapDistance(TF3 aabb_min, TF3 aabb_max, Fn fn) : fn_(std::move(fn)) {
    this->aabb_min_ = aabb_min;
    this->aabb_max_ = aabb_max;
  }
  static MapDistance* create(TF3 aabb_min, TF3 aabb_max, Fn fn) {
    return new MapDistance(aabb_min, aabb_max, std::move(fn));
  }
  alv_distance to_c() const override {
    alv_distance d{};
    d.type = ALV_DISTANCE_MAP;
    d.aabb_min[0] = this->aabb_min_.x;
    d.aabb_min[1] = this->aabb_min_.y;
    d.aabb_min[2] = this->aabb_min_.z;
    d.aabb_max[0] = this->aabb_max_.x;
    d.aabb_max[1] = this->aabb_max_.y;
    d.aabb_max[2] = this->aabb_max_.z;
    d.signed_distance = &MapDistance::call;
    d.context = const_cast<MapDistance*>(this);
    return d;
  }
  Fn fn_;

 private:
  static void call(void* self, const float* points, unsigned n, float* out) {
    static_cast<MapDistance*>(self)->fn_(points, n, out);
  }
};
}  // namespace dg

enum class VolumeMethod { volume_map = 0, pellets = 1 };

// Mesh (mesh.hpp:13-42), the part rigid contacts use: the vertex list of a
// collision mesh and the vertex generators of mesh.cu (set_box :206-276,
// set_uv_sphere :11-63, translate / rotate / scale :278-294).  Faces,
// normals and texture coordinates serve the reference's display only.
struct Mesh {
  std::vector<float3> vertices;
  void clear() { vertices.clear(); }
  // n x n lattice on each face of a box of the given widths, faces in the
  // order +z, -z, +x, -x, -y, +y
  void set_box(float3 widths, U n = 2) {
    clear();
    const float step = 1.0f / float(n - 1);
    for (U face = 0; face < 6; ++face)
      for (U j = 0; j < n; ++j)
        for (U k = 0; k < n; ++k) {
          const float a = -0.5f + step * float(k), b = -0.5f + step * float(j);
          float3 c;
          switch (face) {
            case 0: c = float3{a, b, 0.5f}; break;
            case 1: c = float3{a, b, -0.5f}; break;
            case 2: c = float3{0.5f, a, b}; break;
            case 3: c = float3{-0.5f, a, b}; break;
            case 4: c = float3{a, -0.5f, b}; break;
            default: c = float3{a, 0.5f, b}; break;
          }
          vertices.push_back(float3{c.x * widths.x, c.y * widths.y, c.z * widths.z});
        }
  }
  // (num_stacks + 1) rings from the +z pole to the -z pole, each with
  // num_sectors + 1 vertices (the first repeated at the end)
  void set_uv_sphere(float radius, U num_sectors, U num_stacks) {
    clear();
    const float pi = 3.14159265358979323846f;
    const float sector_step = 2 * pi / float(num_sectors);
    const float stack_step = pi / float(num_stacks);
    for (U i = 0; i <= num_stacks; ++i) {
      const float stack_angle = pi / 2 - float(i) * stack_step;
      const float xy = radius * std::cos(stack_angle), z = radius * std::sin(stack_angle);
      for (U j = 0; j <= num_sectors; ++j) {
        const float sector_angle = float(j) * sector_step;
        vertices.push_back(float3{xy * std::cos(sector_angle), xy * std::sin(sector_angle), z});
      }
    }
  }
  void translate(float3 dx) {
    for (float3& v : vertices) v = float3{v.x + dx.x, v.y + dx.y, v.z + dx.z};
  }
  void rotate(float4 q) {
    for (float3& v : vertices) {
      const float x = (1 - 2 * (q.y * q.y + q.z * q.z)) * v.x + 2 * (q.x * q.y - q.z * q.w) * v.y +
                      2 * (q.x * q.z + q.y * q.w) * v.z;
      const float y = 2 * (q.x * q.y + q.z * q.w) * v.x + (1 - 2 * (q.x * q.x + q.z * q.z)) * v.y +
                      2 * (q.y * q.z - q.x * q.w) * v.z;
      const float z = 2 * (q.x * q.z - q.y * q.w) * v.x + 2 * (q.y * q.z + q.x * q.w) * v.y +
                      (1 - 2 * (q.x * q.x + q.y * q.y)) * v.z;
      v = float3{x, y, z};
    }
  }
  void scale(float s) {
    for (float3& v : vertices) v = float3{v.x * s, v.y * s, v.z * s};
  }
};

// Contact<float3, float> (contact.hpp:7-24)
using Contact = alv_contact;

// Pile<float> (pile.hpp:20-1066): boundaries with volume maps and rigid
// kinematics.  Host kinematics are exposed through get/set helpers.
template <typename TF>
class Pile {
 public:
  Pile(Store& store, Runner<TF>& runner, U max_num_contacts,
       VolumeMethod volume_method = VolumeMethod::volume_map, U max_num_pellets = 10000,
       Const<TF>* cn = nullptr, ConstiN* cni = nullptr)
      : store_(store), cn_(cn ? cn : &store.get_cn()), cni_(cni ? cni : &store.get_cni()) {
    (void)runner;
    check(alv_pile_create_ex(&pile_, max_num_contacts, static_cast<int>(volume_method),
                             max_num_pellets));
    max_num_contacts_ = max_num_contacts;
    cni_->max_num_contacts = max_num_contacts;
    if (cn)
      store.copy_cn_external(*cn_, *cni_);
    else
      store.copy_cn();
  }
  ~Pile() { alv_pile_destroy(pile_); }
  Pile(Pile const&) = delete;
  Pile& operator=(Pile const&) = delete;

  U add(dg::Distance<float3, TF>* distance, U3 resolution = U3{0, 0, 0}, TF cell_width = 0,
        TF margin = std::numeric_limits<TF>::infinity(), TF sign = 1,
        Mesh const& collision_mesh = Mesh(), TF mass = 0, TF restitution = 1, TF friction = 0,
        float3 inertia_tensor = float3{1, 1, 1}, float3 x = float3{0, 0, 0},
        float4 q = float4{0, 0, 0, 1}, Mesh const& display_mesh = Mesh()) {
    (void)display_mesh;
    distances_.emplace_back(distance);
    alv_distance d = distance->to_c();
    alv_const cn = cn_->to_c();
    unsigned res[3] = {resolution.x, resolution.y, resolution.z};
    float inertia[3] = {inertia_tensor.x, inertia_tensor.y, inertia_tensor.z};
    float xx[3] = {x.x, x.y, x.z};
    float qq[4] = {q.x, q.y, q.z, q.w};
    unsigned id = 0;
    check(alv_pile_add(pile_, &cn, &d, res, cell_width, margin, sign, mass, restitution,
                       friction, inertia, xx, qq, &id));
    if (!collision_mesh.vertices.empty())
      check(alv_pile_set_collision_vertices(pile_, id,
                                            reinterpret_cast<const float*>(collision_mesh.vertices.data()),
                                            U(collision_mesh.vertices.size())));
    mass_.push_back(mass);
    restitution_.push_back(restitution);
    friction_.push_back(friction);
    inertia_tensor_.push_back(inertia_tensor);
    cni_->num_boundaries = get_size();
    return id;
  }
  // Pile::add with pellets (pile.hpp:217-283): the pellets (body frame) are
  // the body's collision vertices; a pellet pile builds no volume map.
  U add_pellets(dg::Distance<float3, TF>* distance, U3 resolution, TF cell_width, TF margin,
                TF sign, std::vector<float3> const& pellets, TF mass = 0, TF restitution = 1,
                TF friction = 0, float3 inertia_tensor = float3{1, 1, 1},
                float3 x = float3{0, 0, 0}, float4 q = float4{0, 0, 0, 1}) {
    Mesh m;
    m.vertices = pellets;
    return add(distance, resolution, cell_width, margin, sign, m, mass, restitution, friction,
               inertia_tensor, x, q);
  }
  // num_pellets_ / max_num_pellets_ (pile.hpp:124-130)
  U num_pellets() const {
    unsigned n = 0, mx = 0;
    check(alv_pile_num_pellets(pile_, &n, &mx));
    return n;
  }
  U max_num_pellets() const {
    unsigned n = 0, mx = 0;
    check(alv_pile_num_pellets(pile_, &n, &mx));
    return mx;
  }
  void hint_identical_sequence(U begin_id, U end_id) {
    check(alv_pile_hint_identical_sequence(pile_, begin_id, end_id));
  }
  // Pile::find_contacts(particle_x, max_num_beads) (pile.hpp:858-904)
  U find_contacts(Variable<1, float3> const& particle_x, U max_num_beads) {
    unsigned n = 0;
    check(alv_pile_find_contacts_pellets(pile_, static_cast<const float*>(particle_x.ptr_),
                                         max_num_beads, cn_->contact_tolerance, &n));
    return n;
  }
  U get_size() const {
    unsigned n = 0;
    check(alv_pile_get_size(pile_, &n));
    return n;
  }
  float3 get_x(U i) const { return kin(i).x; }
  float3 get_v(U i) const { return kin(i).v; }
  float4 get_q(U i) const { return kin(i).q; }
  float3 get_omega(U i) const { return kin(i).omega; }
  float3 get_force(U i) const { return kin(i).force; }
  float3 get_torque(U i) const { return kin(i).torque; }
  void set_x(U i, float3 x) { check(alv_pile_set_kinematics(pile_, i, &x.x, nullptr, nullptr, nullptr)); }
  void set_v(U i, float3 v) { check(alv_pile_set_kinematics(pile_, i, nullptr, &v.x, nullptr, nullptr)); }
  void set_q(U i, float4 q) { check(alv_pile_set_kinematics(pile_, i, nullptr, nullptr, &q.x, nullptr)); }
  void set_omega(U i, float3 w) {
    check(alv_pile_set_kinematics(pile_, i, nullptr, nullptr, nullptr, &w.x));
  }
  // mass_, restitution_, friction_, inertia_tensor_ of body i (pile.hpp:97-101)
  struct Body {
    TF mass, restitution, friction;
    float3 inertia_tensor;
  };
  Body get_body(U i) const {
    Body b;
    check(alv_pile_get_body(pile_, i, &b.mass, &b.restitution, &b.friction, &b.inertia_tensor.x));
    return b;
  }
  void set_body(U i, Body const& b) {
    check(alv_pile_set_body(pile_, i, b.mass, b.restitution, b.friction, &b.inertia_tensor.x));
  }
  // Pile::replace (pile.hpp:423-492)
  void replace(U i, dg::Distance<float3, TF>* distance, U3 resolution = U3{0, 0, 0},
               TF cell_width = 0, TF margin = std::numeric_limits<TF>::infinity(), TF sign = 1,
               Mesh const& collision_mesh = Mesh(), TF mass = 0, TF restitution = 1,
               TF friction = 0, float3 inertia_tensor = float3{1, 1, 1},
               float3 x = float3{0, 0, 0}, float4 q = float4{0, 0, 0, 1},
               Mesh const& display_mesh = Mesh()) {
    (void)display_mesh;
    if (i >= distances_.size()) throw std::out_of_range("no boundary " + std::to_string(i));
    alv_distance d = distance->to_c();
    alv_const cn = cn_->to_c();
    unsigned res[3] = {resolution.x, resolution.y, resolution.z};
    float inertia[3] = {inertia_tensor.x, inertia_tensor.y, inertia_tensor.z};
    float xx[3] = {x.x, x.y, x.z};
    float qq[4] = {q.x, q.y, q.z, q.w};
    check(alv_pile_replace(pile_, i, &cn, &d, res, cell_width, margin, sign, mass, restitution,
                           friction, inertia, xx, qq));
    distances_[i].reset(distance);
    check(alv_pile_set_collision_vertices(
        pile_, i, reinterpret_cast<const float*>(collision_mesh.vertices.data()),
        U(collision_mesh.vertices.size())));
    mass_[i] = mass;
    restitution_[i] = restitution;
    friction_[i] = friction;
    inertia_tensor_[i] = inertia_tensor;
  }
  // Pile::calculate_cfl_v2 (pile.hpp:816-825)
  TF calculate_cfl_v2() const {
    float v2 = 0;
    check(alv_pile_calculate_cfl_v2(pile_, &v2));
    return v2;
  }
  // Pile::compute_mask (pile.hpp:1048-1065)
  void compute_mask(U i, TF distance_threshold, Variable<1, float3> const& sample_x,
                    Variable<1, TF>& mask, U num_samples) {
    if (sample_x.get_linear_shape() < num_samples || mask.get_linear_shape() < num_samples)
      throw std::out_of_range("sample_x and mask must hold num_samples elements");
    check(alv_pile_compute_mask(pile_, i, distance_threshold,
                                static_cast<const float*>(sample_x.ptr_),
                                static_cast<float*>(mask.ptr_), num_samples));
  }
  // Pile::get_matrix (pile.hpp:1010-1031): the body's rigid transform,
  // column-major like the reference's glm::mat4 (m[12..14] = x).
  std::array<float, 16> get_matrix(U i) const {
    const float4 q = get_q(i);
    const float3 t = get_x(i);
    return {1 - 2 * (q.y * q.y + q.z * q.z), 2 * (q.x * q.y + q.z * q.w),
            2 * (q.x * q.z - q.y * q.w),     0,
            2 * (q.x * q.y - q.z * q.w),     1 - 2 * (q.x * q.x + q.z * q.z),
            2 * (q.y * q.z + q.x * q.w),     0,
            2 * (q.x * q.z + q.y * q.w),     2 * (q.y * q.z - q.x * q.w),
            1 - 2 * (q.x * q.x + q.y * q.y), 0,
            t.x,                             t.y,
            t.z,                             1};
  }
  void set_gravity(float3 g) { check(alv_pile_set_gravity(pile_, &g.x)); }
  void integrate_kinematics(TF dt) { check(alv_pile_integrate_kinematics(pile_, dt)); }
  void copy_kinematics_to_device() { check(alv_pile_copy_kinematics_to_device(pile_)); }
  // Device copies of the kinematics (x_device_, v_device_, q_device_,
  // omega_device_ in the reference, pile.hpp:103-110), as of the last solver
  // step start or copy_kinematics_to_device.
  Variable<1, float3> x_device() { return Variable<1, float3>({get_size()}, kin_dev(0)); }
  Variable<1, float3> v_device() { return Variable<1, float3>({get_size()}, kin_dev(1)); }
  Variable<1, float4> q_device() { return Variable<1, float4>({get_size()}, kin_dev(2)); }
  Variable<1, float3> omega_device() { return Variable<1, float3>({get_size()}, kin_dev(3)); }

  // Rigid contacts (pile.hpp:826-979).  find_contacts() over every pair is
  // the overload the solvers call after integrating the bodies.
  U find_contacts(U i, U j) { return find(int(i), int(j)); }
  U find_contacts(U i) { return find(int(i), -1); }
  U find_contacts() { return find(-1, -1); }
  void solve_contacts() { check(alv_pile_solve_contacts(pile_)); }
  std::vector<Contact> get_contacts() const {
    std::vector<Contact> out(max_num_contacts_);
    unsigned n = 0;
    check(alv_pile_get_contacts(pile_, out.data(), U(out.size()), &n));
    out.resize(n);
    return out;
  }
  void set_collision_vertices(U i, std::vector<float3> const& vertices) {
    check(alv_pile_set_collision_vertices(pile_, i, reinterpret_cast<const float*>(vertices.data()),
                                          U(vertices.size())));
  }

  // Rigid kinematics files (pile.hpp:757-791): per body x*x_scale,
  // v*v_scale (float3 each), q (float4), omega*omega_scale (float3), 52
  // bytes a body, no header.
  static constexpr size_t kKinRecord = sizeof(float) * 13;
  void write_file(const char* filename, TF x_scale = 1, TF v_scale = 1, TF omega_scale = 1) const {
    std::ofstream stream(filename, std::ios::binary | std::ios::trunc);
    if (!stream.is_open()) throw std::runtime_error(std::string("Failed to open file: ") + filename);
    for (U i = 0; i < get_size(); ++i) {
      const Kin k = kin(i);
      const float rec[13] = {k.x.x * x_scale, k.x.y * x_scale, k.x.z * x_scale,
                             k.v.x * v_scale, k.v.y * v_scale, k.v.z * v_scale,
                             k.q.x, k.q.y, k.q.z, k.q.w,
                             k.omega.x * omega_scale, k.omega.y * omega_scale,
                             k.omega.z * omega_scale};
      stream.write(reinterpret_cast<const char*>(rec), kKinRecord);
    }
    if (!stream) throw std::runtime_error(std::string("Failed to write file: ") + filename);
  }
  // Reads up to `num_rigids` records (all bodies when negative), skipping
  // `src_offset` records of the file, into bodies dst_offset, dst_offset+1,
  // ...; stops at the end of the file.  Returns the number of bodies set.
  U read_file(const char* filename, int num_rigids = -1, U dst_offset = 0, U src_offset = 0) {
    std::ifstream stream(filename, std::ios::binary);
    if (!stream.is_open()) throw std::runtime_error(std::string("Failed to open file: ") + filename);
    const U size = get_size();
    U count = num_rigids >= 0 ? std::min(size, U(num_rigids)) : size;
    if (dst_offset >= size) return 0;
    count = std::min(count, size - dst_offset);
    stream.seekg(std::streamoff(kKinRecord) * src_offset, std::ios_base::beg);
    U done = 0;
    for (; done < count; ++done) {
      float rec[13];
      stream.read(reinterpret_cast<char*>(rec), kKinRecord);
      if (!stream) break;
      check(alv_pile_set_kinematics(pile_, dst_offset + done, rec, rec + 3, rec + 6, rec + 10));
    }
    return done;
  }
  static U get_size_from_file(const char* filename) {
    std::ifstream stream(filename, std::ios::binary | std::ios::ate);
    if (!stream.is_open()) throw std::runtime_error(std::string("Failed to open file: ") + filename);
    return U(size_t(stream.tellg()) / kKinRecord);
  }
  Variable<1, float> get_volume_grid(U i) const {
    float* p = nullptr;
    unsigned n = 0;
    check(alv_pile_get_volume_grid(pile_, i, &p, &n));
    return Variable<1, float>({n}, p);
  }
  // Pile::get_distance_grids (pile.hpp:981-988): empty for analytic distances
  Variable<1, float> get_distance_grid(U i) const {
    float* p = nullptr;
    unsigned n = 0;
    check(alv_pile_get_distance_grid(pile_, i, &p, &n));
    return Variable<1, float>({n}, p);
  }
  std::vector<Variable<1, float>> get_distance_grids() const {
    std::vector<Variable<1, float>> out;
    for (U i = 0; i < get_size(); ++i) out.push_back(get_distance_grid(i));
    return out;
  }
  struct GridInfo {
    float3 domain_min, domain_max, cell_size;
    U3 resolution;
    float sign, mass;
  };
  GridInfo get_grid_info(U i) const {
    GridInfo g;
    check(alv_pile_get_grid_info(pile_, i, &g.domain_min.x, &g.domain_max.x, &g.resolution.x,
                                 &g.cell_size.x, &g.sign, &g.mass));
    return g;
  }
  alv_pile* handle() { return pile_; }

  std::vector<TF> mass_, restitution_, friction_;
  std::vector<float3> inertia_tensor_;

 private:
  struct Kin {
    float3 x, v, omega, force, torque;
    float4 q;
  };
  float* kin_dev(int which) {
    float* p[4] = {nullptr, nullptr, nullptr, nullptr};
    check(alv_pile_kinematics_device(pile_, &p[0], &p[1], &p[2], &p[3]));
    return p[which];
  }
  U find(int i, int j) {
    unsigned n = 0;
    check(alv_pile_find_contacts(pile_, i, j, cn_->contact_tolerance, &n));
    return n;
  }
  Kin kin(U i) const {
    Kin k;
    check(alv_pile_get_kinematics(pile_, i, &k.x.x, &k.v.x, &k.q.x, &k.omega.x, &k.force.x,
                                  &k.torque.x));
    return k;
  }
  Store& store_;
  Const<TF>* cn_;
  ConstiN* cni_;
  alv_pile* pile_ = nullptr;
  U max_num_contacts_ = 0;
  std::vector<std::unique_ptr<dg::Distance<float3, TF>>> distances_;
};

// Learned warm-start network (no counterpart in the reference): owns an
// alv_mlp.  Data-set arguments are device pointers with their element
// counts (pairs), features [n][8] and targets [n][2] in FP32.
class WarmStartNet {
 public:
  explicit WarmStartNet(unsigned long long seed = 0) { check(alv_mlp_create(&net_, seed)); }
  ~WarmStartNet() { alv_mlp_destroy(net_); }
  WarmStartNet(const WarmStartNet&) = delete;
  WarmStartNet& operator=(const WarmStartNet&) = delete;
  static int num_params() { return alv_mlp_num_params(); }
  std::vector<float> get_params() const {
    std::vector<float> p(static_cast<size_t>(num_params()));
    check(alv_mlp_get_params(net_, p.data()));
    return p;
  }
  void set_params(std::vector<float> const& p) {
    if (p.size() != size_t(num_params())) throw std::invalid_argument("expected 9026 parameters");
    check(alv_mlp_set_params(net_, p.data()));
  }
  std::vector<float> get_normalization() const {
    std::vector<float> n(20);
    check(alv_mlp_get_normalization(net_, n.data()));
    return n;
  }
  void set_normalization(std::vector<float> const& n) {
    if (n.size() != 20) throw std::invalid_argument("expected 20 normalisation values");
    check(alv_mlp_set_normalization(net_, n.data()));
  }
  void fit_normalization(const float* x, const float* y, size_t n) {
    check(alv_mlp_fit_normalization(net_, x, y, n));
  }
  void forward(const float* x, size_t n, float* y) { check(alv_mlp_forward(net_, x, n, y)); }
  double loss_grad(const float* xn, const float* tn, size_t n, std::vector<float>& grad) {
    grad.resize(size_t(num_params()));
    double loss = 0;
    check(alv_mlp_loss_grad(net_, xn, tn, n, &loss, grad.data()));
    return loss;
  }
  std::vector<float> train(const float* x, const float* y, size_t n, U batch, U steps, float lr,
                           unsigned long long seed) {
    std::vector<float> loss(steps);
    check(alv_mlp_train(net_, x, y, n, batch, steps, lr, seed, loss.data()));
    return loss;
  }
  alv_mlp* handle() const { return net_; }

 private:
  alv_mlp* net_ = nullptr;
};

// One scene of a batch: its pile of static boundaries, particle count and
// neighbour grid (no counterpart in the reference).
template <typename TF>
struct SceneSpec {
  Pile<TF>* pile;
  U num_particles;
  U3 grid_res;
  I3 grid_offset;
};

// Solver base (solver.hpp:8-306) over alv_solver.
template <typename TF>
class SolverBase {
 public:
  // a batch of independent scenes in one solver
  SolverBase(int kind, Store& store, std::vector<SceneSpec<TF>> const& scenes,
             Const<TF> const* cn, ConstiN const* cni)
      : max_num_particles(batch_capacity(scenes)), store_(store), pile_(*scenes.at(0).pile) {
    if (cn)
      store.copy_cn_external(*cn, *cni);
    else
      store.copy_cn();
    alv_const c = store.device_cn().to_c();
    alv_consti ci = store.device_cni().to_c();
    std::vector<alv_scene_desc> d(scenes.size());
    for (size_t k = 0; k < scenes.size(); ++k) {
      d[k].pile = scenes[k].pile->handle();
      d[k].num_particles = scenes[k].num_particles;
      d[k].grid_res[0] = scenes[k].grid_res.x;
      d[k].grid_res[1] = scenes[k].grid_res.y;
      d[k].grid_res[2] = scenes[k].grid_res.z;
      d[k].grid_offset[0] = scenes[k].grid_offset.x;
      d[k].grid_offset[1] = scenes[k].grid_offset.y;
      d[k].grid_offset[2] = scenes[k].grid_offset.z;
    }
    check(alv_solver_create_batch(&solver_, kind, U(d.size()), d.data(), &c, &ci));
    check(alv_solver_get_settings(solver_, &set_));
    particle_radius = store.device_cn().particle_radius;
  }
  static U batch_capacity(std::vector<SceneSpec<TF>> const& scenes) {
    U n = 0;
    for (auto const& s : scenes) n += (s.num_particles + 255u) / 256u * 256u;
    return n;
  }
  U num_scenes() const {
    U n = 0;
    check(alv_solver_num_scenes(solver_, &n));
    return n;
  }
  std::pair<U, U> scene_range(U k) const {
    U first = 0, count = 0;
    check(alv_solver_scene_range(solver_, k, &first, &count));
    return {first, count};
  }
  alv_solver_report scene_report(U k) const {
    alv_solver_report r;
    check(alv_solver_scene_report(solver_, k, &r));
    return r;
  }

  SolverBase(int kind, Runner<TF>& runner, Pile<TF>& pile, Store& store, U max_num_particles,
             U num_ushers, bool enable_surface_tension, bool enable_vorticity,
             Const<TF> const* cn, ConstiN const* cni, bool graphical)
      : max_num_particles(max_num_particles), store_(store), pile_(pile) {
    (void)runner;
    if (num_ushers) throw std::runtime_error("ushers are not supported on this path");
    if (enable_surface_tension || enable_vorticity)
      throw std::runtime_error("surface tension and vorticity are not supported on this path");
    if (graphical) throw std::runtime_error("graphical variables are not supported");
    if (cn)
      store.copy_cn_external(*cn, *cni);
    else
      store.copy_cn();
    alv_const c = store.device_cn().to_c();
    alv_consti ci = store.device_cni().to_c();
    check(alv_solver_create(&solver_, kind, pile.handle(), max_num_particles, &c, &ci));
    check(alv_solver_get_settings(solver_, &set_));
    particle_radius = store.device_cn().particle_radius;
  }
  virtual ~SolverBase() { alv_solver_destroy(solver_); }
  SolverBase(SolverBase const&) = delete;
  SolverBase& operator=(SolverBase const&) = delete;

  // settings (read-write fields of the reference)
  U get_num_particles() const { return set_.num_particles; }
  void set_num_particles(U n) { set_.num_particles = n; push(); }
#define ALV_RW(name)                                 \
  TF get_##name() const { return set_.name; }        \
  void set_##name(TF value) {                        \
    set_.name = value;                               \
    push();                                          \
  }
  ALV_RW(t)
  ALV_RW(dt)
  ALV_RW(initial_dt)
  ALV_RW(max_dt)
  ALV_RW(min_dt)
  ALV_RW(cfl)
#undef ALV_RW
  alv_solver_report report() const {
    alv_solver_report r;
    check(alv_solver_get_report(solver_, &r));
    return r;
  }
  void reset_solving_var() { check(alv_solver_reset_solving_var(solver_)); }
  void reset_t() {
    check(alv_solver_reset_t(solver_));
    pull();
  }
  template <U wrap = 0>
  void update_particle_neighbors() {
    sync_const();
    push();
    check(alv_solver_update_particle_neighbors(solver_));
  }
  // Solver::transform_all_pellets (solver.hpp:204-223)
  void transform_all_pellets() {
    sync_const();
    push();
    check(alv_solver_transform_all_pellets(solver_));
  }
  void compute_all_boundaries() {}
  // step execution as replayed HIP graphs (no counterpart in the reference)
  void set_graph(bool enable, U slack = 0) { check(alv_solver_set_graph(solver_, enable, slack)); }
  void set_tiles(bool enable) { check(alv_solver_set_tiles(solver_, enable)); }
  // learned warm start (no counterpart in the reference); null = plain
  void set_warm_start_net(WarmStartNet const* net, bool divergence = true, bool density = true) {
    check(alv_solver_set_warm_start_net(solver_, net ? net->handle() : nullptr, divergence,
                                        density));
  }
  void set_feature_log(bool enable) { check(alv_solver_set_feature_log(solver_, enable)); }
  // training pairs of the last step into device buffers of at least
  // num_particles pairs
  void export_training_pairs(float* x, float* y) {
    check(alv_solver_export_training_pairs(solver_, x, y));
  }
  // compute_velocity_of_density_change's density-change sum alone
  // (runner.hpp:3297-3314) into particle_density_adv
  void compute_density_change() {
    sync_const();
    push();
    check(alv_solver_compute_density_change(solver_));
  }
  // integrate_velocity<wrap> (runner.hpp:3613-3623)
  void integrate_velocity() {
    sync_const();
    push();
    check(alv_solver_integrate_velocity(solver_));
  }
  template <U wrap = 0>
  void step() {
    sync_const();
    push();
    check(alv_solver_step(solver_));
    pull();
  }
  template <typename M>
  Variable<1, M> field1(int f) {
    void* p = nullptr;
    size_t n = 0;
    unsigned c = 0;
    check(alv_solver_get_field(solver_, f, &p, &n, &c));
    return Variable<1, M>({U(n)}, p);
  }
  Variable<1, float3> particle_x() { return field1<float3>(ALV_FIELD_X); }
  Variable<1, float3> particle_v() { return field1<float3>(ALV_FIELD_V); }
  Variable<1, float3> particle_a() { return field1<float3>(ALV_FIELD_A); }
  Variable<1, float> particle_density() { return field1<float>(ALV_FIELD_DENSITY); }
  Variable<1, U> particle_num_neighbors() { return field1<U>(ALV_FIELD_NUM_NEIGHBORS); }
  Variable<1, U> grid_anomaly() { return field1<U>(ALV_FIELD_GRID_ANOMALY); }
  Variable<1, U> boundary_mask() { return field1<U>(ALV_FIELD_BOUNDARY_MASK); }
  // [rows][n / rows] view of a reference-layout field (the rows are
  // boundaries, or particles for the neighbour list)
  template <typename M>
  Variable<2, M> field2(int f, bool rows_are_particles) {
    void* p = nullptr;
    size_t n = 0;
    unsigned c = 0;
    check(alv_solver_get_field(solver_, f, &p, &n, &c));
    U rows = rows_are_particles ? max_num_particles : U(n / max_num_particles);
    return Variable<2, M>({rows, U(n / rows)}, p);
  }
  // particle_neighbors (solver.hpp:74-79): [max_num_particles][K] float4
  // (x_i - x_j, bits of j), made from the last neighbour search on request
  Variable<2, float4> particle_neighbors() { return field2<float4>(ALV_FIELD_NEIGHBORS, true); }
  // particle_boundary, particle_boundary_kernel, particle_force,
  // particle_torque (solver.hpp:39-46): [num_boundaries][max_num_particles]
  Variable<2, float4> particle_boundary() { return field2<float4>(ALV_FIELD_BOUNDARY, false); }
  Variable<2, float4> particle_boundary_kernel() {
    return field2<float4>(ALV_FIELD_BOUNDARY_KERNEL, false);
  }
  Variable<2, float3> particle_force() { return field2<float3>(ALV_FIELD_FORCE, false); }
  Variable<2, float3> particle_torque() { return field2<float3>(ALV_FIELD_TORQUE, false); }
  // static boundaries accumulate particle_force / particle_torque too
  void set_static_boundary_forces(bool enable) {
    check(alv_solver_set_static_boundary_forces(solver_, enable));
  }
  Variable<3, U> pid_length() {
    void* p = nullptr;
    check(alv_solver_get_field(solver_, ALV_FIELD_PID_LENGTH, &p, nullptr, nullptr));
    ConstiN const& ci = store_.device_cni();
    return Variable<3, U>({ci.grid_res.x, ci.grid_res.y, ci.grid_res.z}, p);
  }
  // neighbour ids in the reference layout [N, Kmax] (host vectors)
  void get_neighbors(std::vector<U>& ids, std::vector<U>& counts) {
    ids.assign(size_t(max_num_particles) * store_.device_cni().max_num_neighbors_per_particle,
               0xffffffffu);
    counts.assign(max_num_particles, 0);
    check(alv_solver_get_neighbors(solver_, ids.data(), counts.data()));
  }
  // pellet lists (which 0: pellets near each fluid particle, as particle ids
  // max_num_particles + k; 1: fluid particles near each pellet)
  void get_pellet_neighbors(int which, U rows, std::vector<U>& ids, std::vector<U>& counts) {
    ids.assign(size_t(rows) * store_.device_cni().max_num_neighbors_per_particle, 0xffffffffu);
    counts.assign(rows, 0);
    check(alv_solver_get_pellet_neighbors(solver_, which, ids.data(), counts.data()));
  }
  void get_grid(std::vector<U>& count, std::vector<U>& start, std::vector<U>& anomaly) {
    ConstiN const& ci = store_.device_cni();
    size_t C = size_t(ci.grid_res.x) * ci.grid_res.y * ci.grid_res.z;
    count.assign(C, 0);
    start.assign(C + 1, 0);
    anomaly.assign(3, 0);
    check(alv_solver_get_grid(solver_, count.data(), start.data(), anomaly.data()));
  }
  alv_solver* handle() { return solver_; }
  SolverGrid pid() { return SolverGrid{solver_}; }
  // Solver::sample_all_boundaries (solver.hpp:186-202)
  void sample_all_boundaries(Variable<1, float3> const& sample_x,
                             Variable<2, float4>& sample_boundary,
                             Variable<2, float4>& sample_boundary_kernel, U num_samples) {
    if (sample_x.shape_[0] < num_samples || sample_boundary.shape_[1] != num_samples ||
        sample_boundary_kernel.shape_[1] != num_samples ||
        sample_boundary.shape_[0] < pile_.get_size() ||
        sample_boundary_kernel.shape_[0] < pile_.get_size())
      throw std::out_of_range("sample boundary buffers must be [num_boundaries][num_samples]");
    check(alv_solver_sample_all_boundaries(solver_, static_cast<const float*>(sample_x.ptr_),
                                           num_samples, static_cast<float*>(sample_boundary.ptr_),
                                           static_cast<float*>(sample_boundary_kernel.ptr_)));
  }

  const U max_num_particles;
  TF particle_radius = 0;

 protected:
  void push() { check(alv_solver_set_settings(solver_, &set_)); }
  void pull() { check(alv_solver_get_settings(solver_, &set_)); }
  void sync_const() {
    alv_const c = store_.device_cn().to_c();
    alv_consti ci = store_.device_cni().to_c();
    check(alv_solver_set_const(solver_, &c, &ci));
  }
  Store& store_;
  Pile<TF>& pile_;
  alv_solver* solver_ = nullptr;
  alv_solver_settings set_{};
};

// SolverDf<float> (solver_df.hpp:10-386)
template <typename TF>
class SolverDf : public SolverBase<TF> {
 public:
  SolverDf(Runner<TF>& runner, Pile<TF>& pile, Store& store, U max_num_particles,
           U num_ushers = 0, bool enable_surface_tension = false, bool enable_vorticity = false,
           Const<TF> const* cn = nullptr, ConstiN const* cni = nullptr, bool graphical = false)
      : SolverBase<TF>(ALV_SOLVER_DF, runner, pile, store, max_num_particles, num_ushers,
                       enable_surface_tension, enable_vorticity, cn, cni, graphical) {}
  // batch of independent scenes
  SolverDf(Store& store, std::vector<SceneSpec<TF>> const& scenes, Const<TF> const* cn = nullptr,
           ConstiN const* cni = nullptr)
      : SolverBase<TF>(ALV_SOLVER_DF, store, scenes, cn, cni) {}
#define ALV_RWS(type, name)                            \
  type get_##name() const { return type(this->set_.name); } \
  void set_##name(type value) {                        \
    this->set_.name = value;                           \
    this->push();                                      \
  }
  ALV_RWS(bool, enable_divergence_solve)
  ALV_RWS(bool, enable_density_solve)
  ALV_RWS(TF, density_change_tolerance)
  ALV_RWS(TF, density_error_tolerance)
  ALV_RWS(U, min_density_solve)
  ALV_RWS(U, max_density_solve)
  ALV_RWS(U, min_divergence_solve)
  ALV_RWS(U, max_divergence_solve)
#undef ALV_RWS
  Variable<1, float> particle_dfsph_factor() { return this->template field1<float>(ALV_FIELD_DFSPH_FACTOR); }
  Variable<1, float> particle_kappa() { return this->template field1<float>(ALV_FIELD_KAPPA); }
  Variable<1, float> particle_kappa_v() { return this->template field1<float>(ALV_FIELD_KAPPA_V); }
  Variable<1, float> particle_density_adv() {
    return this->template field1<float>(ALV_FIELD_DENSITY_ADV);
  }
};

// SolverIi<float> (solver_ii.hpp:8-266)
template <typename TF>
class SolverIi : public SolverBase<TF> {
 public:
  SolverIi(Runner<TF>& runner, Pile<TF>& pile, Store& store, U max_num_particles,
           U num_ushers = 0, bool enable_surface_tension = false, bool enable_vorticity = false,
           Const<TF> const* cn = nullptr, ConstiN const* cni = nullptr, bool graphical = false)
      : SolverBase<TF>(ALV_SOLVER_II, runner, pile, store, max_num_particles, num_ushers,
                       enable_surface_tension, enable_vorticity, cn, cni, graphical) {}
  TF get_density_error_tolerance() const { return this->set_.density_error_tolerance; }
  void set_density_error_tolerance(TF v) {
    this->set_.density_error_tolerance = v;
    this->push();
  }
  U get_min_density_solve() const { return this->set_.min_density_solve; }
  void set_min_density_solve(U v) {
    this->set_.min_density_solve = v;
    this->push();
  }
  U get_max_density_solve() const { return this->set_.max_density_solve; }
  void set_max_density_solve(U v) {
    this->set_.max_density_solve = v;
    this->push();
  }
  Variable<1, float> particle_pressure() { return this->template field1<float>(ALV_FIELD_PRESSURE); }
  // the IISPH arrays (solver_ii.hpp:53-62), user order
  Variable<1, float> particle_last_pressure() {
    return this->template field1<float>(ALV_FIELD_LAST_PRESSURE);
  }
  Variable<1, float> particle_aii() { return this->template field1<float>(ALV_FIELD_AII); }
  Variable<1, float3> particle_dii() { return this->template field1<float3>(ALV_FIELD_DII); }
  Variable<1, float3> particle_dij_pj() { return this->template field1<float3>(ALV_FIELD_DIJ_PJ); }
  Variable<1, float> particle_adv_density() {
    return this->template field1<float>(ALV_FIELD_ADV_DENSITY);
  }
  Variable<1, float3> particle_pressure_accel() {
    return this->template field1<float3>(ALV_FIELD_PRESSURE_ACCEL);
  }
  Variable<1, float> particle_density_err() {
    return this->template field1<float>(ALV_FIELD_DENSITY_ERR);
  }

 protected:
  SolverIi(int kind, Runner<TF>& runner, Pile<TF>& pile, Store& store, U max_num_particles,
           U num_ushers, bool enable_surface_tension, bool enable_vorticity,
           Const<TF> const* cn, ConstiN const* cni, bool graphical)
      : SolverBase<TF>(kind, runner, pile, store, max_num_particles, num_ushers,
                       enable_surface_tension, enable_vorticity, cn, cni, graphical) {}
};

// SolverI<float> (solver_i.hpp:8-353): pressure-Laplacian ISPH with a
// weighted Jacobi solve; the same settings and report as SolverIi.
template <typename TF>
class SolverI : public SolverIi<TF> {
 public:
  SolverI(Runner<TF>& runner, Pile<TF>& pile, Store& store, U max_num_particles,
          U num_ushers = 0, bool enable_surface_tension = false, bool enable_vorticity = false,
          Const<TF> const* cn = nullptr, ConstiN const* cni = nullptr, bool graphical = false)
      : SolverIi<TF>(ALV_SOLVER_I, runner, pile, store, max_num_particles, num_ushers,
                     enable_surface_tension, enable_vorticity, cn, cni, graphical) {}
};

}  // namespace alluvion
