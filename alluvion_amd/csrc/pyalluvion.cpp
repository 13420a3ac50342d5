// pyalluvion.cpp -- pybind11 module alluvion_amd._alluvion.
//
// Exposes the reference's Python names for the solver path
// (lib/pyalluvion.cu:1279-1486): Store with create{D}D{type}{n} /
// remove{...}, Variable{D}D{type}{n}, Runnerfloat, Constfloat, ConstiN,
// Pilefloat, SolverDffloat, SolverIifloat, dg.{Box,BoxShell,Sphere}Distancefloat,
// the vector types and the NumericType / VolumeMethod enums.  Everything is
// a thin layer over alv/alluvion.hpp, which itself calls the C ABI (alv.h).
#include <pybind11/numpy.h>
#include <cstring>

#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <tuple>

#include "alv/alluvion.hpp"

namespace py = pybind11;
using namespace alluvion;

namespace {

template <typename V>
void declare_vec(py::module& m, const char* name) {}

// Device pointer of any bound Variable with room for `floats` FP32 values.
float* var_floats(py::object var, size_t floats, const char* what) {
  const size_t bytes = var.attr("get_num_bytes")().cast<size_t>();
  if (bytes < floats * sizeof(float))
    throw std::invalid_argument(std::string(what) + ": variable too small");
  return reinterpret_cast<float*>(var.attr("ptr").cast<uintptr_t>());
}

template <U D, typename M>
void declare_variable(py::module& m, py::class_<Store>& store, const std::string& name) {
  using Var = Variable<D, M>;
  store.def(("create" + name).c_str(),
            [](Store& s, std::vector<U> shape) {
              if (shape.size() != D) throw std::invalid_argument("shape rank mismatch");
              std::array<U, D> sh;
              for (U i = 0; i < D; ++i) sh[i] = shape[i];
              return s.create<D, M>(sh);
            },
            py::return_value_policy::take_ownership);
  store.def(("remove" + name).c_str(), [](Store& s, Var& v) { s.remove<D, M>(v); });
  py::class_<Var>(m, ("Variable" + name).c_str())
      .def(py::init<const Var&>())
      .def_property_readonly("ptr", [](Var const& v) { return reinterpret_cast<uintptr_t>(v.ptr_); })
      .def("get_bytes",
           [](Var& v, py::array_t<unsigned char> bytes, U offset) {
             v.get_bytes(bytes.mutable_data(), U(bytes.size()), offset);
           },
           py::arg("bytes"), py::arg("offset") = 0)
      .def("set_bytes",
           [](Var& v, py::array_t<unsigned char> bytes, U offset) {
             v.set_bytes(bytes.data(), U(bytes.size()), offset);
           },
           py::arg("bytes"), py::arg("offset") = 0)
      .def("set_from", py::overload_cast<Var const&, U, U>(&Var::set_from), py::arg("src"),
           py::arg("num_elements"), py::arg("offset") = 0)
      .def("set_from", py::overload_cast<Var const&>(&Var::set_from), py::arg("src"))
      .def("set_zero", &Var::set_zero)
      .def("set_same", py::overload_cast<int, U, U>(&Var::set_same), py::arg("value"),
           py::arg("num_elements"), py::arg("offset") = 0)
      .def("set_same", py::overload_cast<int>(&Var::set_same), py::arg("value"))
      .def("get_type", &Var::get_type)
      .def("get_num_primitives_per_element", &Var::get_num_primitives_per_element)
      .def("get_linear_shape", &Var::get_linear_shape)
      .def("get_num_primitives", &Var::get_num_primitives)
      .def("get_num_bytes", &Var::get_num_bytes)
      .def("read_file", &Var::read_file, py::arg("filename"))
      .def("write_file", &Var::write_file, py::arg("filename"),
           py::arg("shape_outermost") = 0, py::arg("offset") = 0)
      .def("get_shape", &Var::get_shape);
}

template <typename T>
py::class_<T> vec3_class(py::module& m, const char* name) {
  return py::class_<T>(m, name)
      .def(py::init([](decltype(T::x) x, decltype(T::x) y, decltype(T::x) z) { return T{x, y, z}; }),
           py::arg("x") = 0, py::arg("y") = 0, py::arg("z") = 0)
      .def_readwrite("x", &T::x)
      .def_readwrite("y", &T::y)
      .def_readwrite("z", &T::z)
      .def("__iter__",
           [](T const& v) { return py::iter(py::make_tuple(v.x, v.y, v.z)); })
      .def("__repr__", [name](T const& v) {
        return std::string(name) + "(" + std::to_string(v.x) + ", " + std::to_string(v.y) + ", " +
               std::to_string(v.z) + ")";
      });
}

}  // namespace

PYBIND11_MODULE(_alluvion, m) {
  m.doc() = "MI355X (gfx950) SPH solver core with alluvion's Python names.";

  py::enum_<NumericType>(m, "NumericType")
      .value("f32", NumericType::f32)
      .value("f64", NumericType::f64)
      .value("i32", NumericType::i32)
      .value("u32", NumericType::u32)
      .value("undefined", NumericType::undefined);
  py::enum_<VolumeMethod>(m, "VolumeMethod")
      .value("volume_map", VolumeMethod::volume_map)
      .value("pellets", VolumeMethod::pellets);

  py::class_<float2>(m, "float2")
      .def(py::init([](float x, float y) { return float2{x, y}; }), py::arg("x") = 0,
           py::arg("y") = 0)
      .def_readwrite("x", &float2::x)
      .def_readwrite("y", &float2::y);
  vec3_class<float3>(m, "float3");
  vec3_class<uint3>(m, "uint3");
  vec3_class<int3>(m, "int3");
  py::class_<float4>(m, "float4")
      .def(py::init([](float x, float y, float z, float w) { return float4{x, y, z, w}; }),
           py::arg("x") = 0, py::arg("y") = 0, py::arg("z") = 0, py::arg("w") = 0)
      .def_readwrite("x", &float4::x)
      .def_readwrite("y", &float4::y)
      .def_readwrite("z", &float4::z)
      .def_readwrite("w", &float4::w)
      .def("__iter__",
           [](float4 const& v) { return py::iter(py::make_tuple(v.x, v.y, v.z, v.w)); });

  py::class_<Const<float>>(m, "Constfloat")
      .def(py::init<>())
      .def("set_kernel_radius", &Const<float>::set_kernel_radius)
      .def("set_particle_attr", &Const<float>::set_particle_attr)
      .def("set_wrap_length", &Const<float>::set_wrap_length)
      .def_readonly("kernel_radius", &Const<float>::kernel_radius)
      .def_readonly("kernel_radius_sqr", &Const<float>::kernel_radius_sqr)
      .def_readonly("cubic_kernel_k", &Const<float>::cubic_kernel_k)
      .def_readonly("cubic_kernel_l", &Const<float>::cubic_kernel_l)
      .def_readonly("cubic_kernel_zero", &Const<float>::cubic_kernel_zero)
      .def_readonly("particle_radius", &Const<float>::particle_radius)
      .def_readonly("particle_vol", &Const<float>::particle_vol)
      .def_readonly("particle_mass", &Const<float>::particle_mass)
      .def_readonly("density0", &Const<float>::density0)
      .def_readonly("wrap_length", &Const<float>::wrap_length)
      .def_readonly("wrap_min", &Const<float>::wrap_min)
      .def_readonly("wrap_max", &Const<float>::wrap_max)
      .def_readwrite("wrap_length_x", &Const<float>::wrap_length_x)
      .def_readwrite("wrap_length_z", &Const<float>::wrap_length_z)
      .def_readwrite("viscosity", &Const<float>::viscosity)
      .def_readwrite("boundary_viscosity", &Const<float>::boundary_viscosity)
      .def_readwrite("vorticity_coeff", &Const<float>::vorticity_coeff)
      .def_readwrite("boundary_vorticity_coeff", &Const<float>::boundary_vorticity_coeff)
      .def_readwrite("inertia_inverse", &Const<float>::inertia_inverse)
      .def_readwrite("viscosity_omega", &Const<float>::viscosity_omega)
      .def_readwrite("surface_tension_coeff", &Const<float>::surface_tension_coeff)
      .def_readwrite("surface_tension_boundary_coeff",
                     &Const<float>::surface_tension_boundary_coeff)
      .def_readwrite("gravity", &Const<float>::gravity)
      .def_readwrite("boundary_epsilon", &Const<float>::boundary_epsilon)
      .def_readwrite("dfsph_factor_epsilon", &Const<float>::dfsph_factor_epsilon)
      .def_readwrite("contact_tolerance", &Const<float>::contact_tolerance);

  py::class_<ConstiN>(m, "ConstiN")
      .def(py::init<>())
      .def_readonly("max_num_contacts", &ConstiN::max_num_contacts)
      .def_readwrite("num_boundaries", &ConstiN::num_boundaries)
      .def_readwrite("max_num_particles_per_cell", &ConstiN::max_num_particles_per_cell)
      .def_readwrite("grid_res", &ConstiN::grid_res)
      .def_readwrite("grid_offset", &ConstiN::grid_offset)
      .def_readwrite("max_num_neighbors_per_particle", &ConstiN::max_num_neighbors_per_particle);

  py::class_<Store> store =
      py::class_<Store>(m, "Store")
          .def(py::init<>())
          .def_static("set_device", &Store::set_device)
          .def_static("get_alu_info", &Store::get_alu_info, py::arg("filename"))
          .def("has_display", [](Store const&) { return false; })
          .def_property_readonly("cnfloat", &Store::get_cn, py::return_value_policy::reference_internal)
          .def_property_readonly("cni", &Store::get_cni, py::return_value_policy::reference_internal)
          .def("create_cnfloat", &Store::create_cn)
          .def("copy_cn_externalfloat", &Store::copy_cn_external)
          .def("copy_cnfloat", &Store::copy_cn);

  declare_variable<1, float>(m, store, "1Dfloat");
  declare_variable<1, float2>(m, store, "1Dfloat2");
  declare_variable<1, float3>(m, store, "1Dfloat3");
  declare_variable<1, float4>(m, store, "1Dfloat4");
  declare_variable<2, float>(m, store, "2Dfloat");
  declare_variable<2, float3>(m, store, "2Dfloat3");
  declare_variable<2, float4>(m, store, "2Dfloat4");
  declare_variable<3, float>(m, store, "3Dfloat");
  declare_variable<4, float>(m, store, "4Dfloat");
  declare_variable<1, U>(m, store, "1Duint");
  declare_variable<2, U>(m, store, "2Duint");
  declare_variable<3, U>(m, store, "3Duint");
  declare_variable<1, I>(m, store, "1Dint");

  py::class_<SolverGrid>(m, "SolverGrid");
  using R = Runner<float>;
  auto make_list = [](R&, Variable<1, float3> const& sample_x, SolverGrid pid, py::object pid_length,
                      Variable<2, float4>& nbr, Variable<1, U>& counts, py::object grid_anomaly,
                      U num_samples) {
    (void)pid_length;    // the solver's own grid
    (void)grid_anomaly;  // the solver's grid_anomaly[2] is set on overflow
    R::launch_make_neighbor_list(sample_x, pid, nbr, counts, num_samples);
  };
  auto bead_pellet = [](R&, Variable<1, float3> const& sample_x, SolverGrid pid,
                        py::object pid_length, Variable<2, float4>& bead, Variable<1, U>& nbead,
                        Variable<2, float4>& pel, Variable<1, U>& npel, py::object grid_anomaly,
                        U max_num_beads, U num_samples, U offset) {
    (void)pid_length;
    (void)grid_anomaly;
    R::launch_make_bead_pellet_neighbor_list(sample_x, pid, bead, nbead, pel, npel,
                                             max_num_beads, num_samples, offset);
  };
  py::class_<Runner<float>>(m, "Runnerfloat")
      .def(py::init<>())
      .def("launch_make_bead_pellet_neighbor_list", bead_pellet, py::arg("sample_x"),
           py::arg("pid"), py::arg("pid_length"), py::arg("sample_bead_neighbors"),
           py::arg("sample_num_bead_neighbors"), py::arg("sample_pellet_neighbors"),
           py::arg("sample_num_pellet_neighbors"), py::arg("grid_anomaly"),
           py::arg("max_num_beads"), py::arg("num_samples"), py::arg("offset") = 0)
      .def("launch_make_bead_pellet_neighbor_list_wrap1", bead_pellet, py::arg("sample_x"),
           py::arg("pid"), py::arg("pid_length"), py::arg("sample_bead_neighbors"),
           py::arg("sample_num_bead_neighbors"), py::arg("sample_pellet_neighbors"),
           py::arg("sample_num_pellet_neighbors"), py::arg("grid_anomaly"),
           py::arg("max_num_beads"), py::arg("num_samples"), py::arg("offset") = 0)
      .def("launch_sample_density_with_pellets",
           [](R&, Variable<1, float3> const& sx, Variable<2, float4> const& nbr,
              Variable<1, U> const& cnt, Variable<2, float4> const& pel, Variable<1, U> const& pcnt,
              Variable<1, float>& out, U ns) {
             R::launch_sample_density_with_pellets(sx, nbr, cnt, pel, pcnt, out, ns);
           },
           py::arg("sample_x"), py::arg("sample_neighbors"), py::arg("sample_num_neighbors"),
           py::arg("sample_pellet_neighbors"), py::arg("sample_num_pellet_neighbors"),
           py::arg("sample_density"), py::arg("num_samples"))
      .def("launch_sample_velocity_with_pellets",
           [](R&, Variable<1, float3> const& sx, Variable<1, float3> const& px,
              Variable<1, float> const& pd, Variable<1, float3> const& pv,
              Variable<2, float4> const& nbr, Variable<1, U> const& cnt, Variable<1, float3>& sv,
              Variable<2, float4> const& pel, Variable<1, U> const& pcnt, U ns) {
             R::launch_sample_velocity_with_pellets(sx, px, pd, pv, nbr, cnt, sv, pel, pcnt, ns);
           },
           py::arg("sample_x"), py::arg("particle_x"), py::arg("particle_density"),
           py::arg("particle_v"), py::arg("sample_neighbors"), py::arg("sample_num_neighbors"),
           py::arg("sample_v"), py::arg("sample_pellet_neighbors"),
           py::arg("sample_num_pellet_neighbors"), py::arg("num_samples"))
      .def("launch_sample_vorticity_with_pellets",
           [](R&, Variable<1, float3> const& sx, Variable<1, float3> const& px,
              Variable<1, float> const& pd, Variable<1, float3> const& pv,
              Variable<2, float4> const& nbr, Variable<1, U> const& cnt,
              Variable<1, float3> const& sv, Variable<1, float3>& vort,
              Variable<2, float4> const& pel, Variable<1, U> const& pcnt, U ns) {
             R::launch_sample_vorticity_with_pellets(sx, px, pd, pv, nbr, cnt, sv, vort, pel, pcnt,
                                                     ns);
           },
           py::arg("sample_x"), py::arg("particle_x"), py::arg("particle_density"),
           py::arg("particle_v"), py::arg("sample_neighbors"), py::arg("sample_num_neighbors"),
           py::arg("sample_v"), py::arg("sample_vorticity"), py::arg("sample_pellet_neighbors"),
           py::arg("sample_num_pellet_neighbors"), py::arg("num_samples"))
      .def("launch_compute_particle_boundary_with_pellets",
           [](R&, Variable<1, float4>& out, Variable<2, float4> const& pel,
              Variable<1, U> const& npel, U n) {
             R::launch_compute_particle_boundary_with_pellets(out, pel, npel, n);
           },
           py::arg("particle_boundary_kernel_combined"), py::arg("sample_pellet_neighbors"),
           py::arg("sample_num_pellet_neighbors"), py::arg("num_particles"))
      .def("launch_compute_density_with_pellets",
           [](R&, Variable<2, float4> const& nbr, Variable<1, U> const& cnt,
              Variable<2, float4> const& pel, Variable<1, U> const& npel, Variable<1, float>& out,
              U n) { R::launch_compute_density_with_pellets(nbr, cnt, pel, npel, out, n); },
           py::arg("particle_neighbors"), py::arg("particle_num_neighbors"),
           py::arg("particle_boundary_neighbors"), py::arg("particle_num_boundary_neighbors"),
           py::arg("particle_density"), py::arg("num_particles"))
      .def("launch_compute_dfsph_factor_with_pellets",
           [](R&, Variable<2, float4> const& nbr, Variable<1, U> const& cnt,
              Variable<2, float4> const& pel, Variable<1, U> const& npel, Variable<1, float>& out,
              U n) { R::launch_compute_dfsph_factor_with_pellets(nbr, cnt, pel, npel, out, n); },
           py::arg("particle_neighbors"), py::arg("particle_num_neighbors"),
           py::arg("particle_boundary_neighbors"), py::arg("particle_num_boundary_neighbors"),
           py::arg("particle_dfsph_factor"), py::arg("num_particles"))
      .def("launch_make_neighbor_list", make_list, py::arg("sample_x"), py::arg("pid"),
           py::arg("pid_length"), py::arg("sample_neighbors"), py::arg("sample_num_neighbors"),
           py::arg("grid_anomaly"), py::arg("num_samples"))
      .def("launch_make_neighbor_list_wrap1", make_list, py::arg("sample_x"), py::arg("pid"),
           py::arg("pid_length"), py::arg("sample_neighbors"), py::arg("sample_num_neighbors"),
           py::arg("grid_anomaly"), py::arg("num_samples"))
      .def("launch_sample_fluid", [](R&, Variable<1, float3> const& sx, Variable<1, float3> const& px,
                                     Variable<1, float> const& pd, Variable<1, float> const& q,
                                     Variable<2, float4> const& nbr, Variable<1, U> const& cnt,
                                     Variable<1, float>& out, U ns) {
             R::launch_sample_fluid(sx, px, pd, q, nbr, cnt, out, ns);
           })
      .def("launch_sample_fluid", [](R&, Variable<1, float3> const& sx, Variable<1, float3> const& px,
                                     Variable<1, float> const& pd, Variable<1, float3> const& q,
                                     Variable<2, float4> const& nbr, Variable<1, U> const& cnt,
                                     Variable<1, float3>& out, U ns) {
             R::launch_sample_fluid(sx, px, pd, q, nbr, cnt, out, ns);
           })
      .def("launch_sample_fluid_density",
           [](R&, Variable<2, float4> const& nbr, Variable<1, U> const& cnt,
              Variable<1, float>& out, U ns) { R::launch_sample_fluid_density(nbr, cnt, out, ns); })
      .def("launch_sample_density",
           [](R&, Variable<1, float3> const& sx, Variable<2, float4> const& nbr,
              Variable<1, U> const& cnt, Variable<1, float>& out, Variable<2, float4> const& ker,
              U ns) { R::launch_sample_density(sx, nbr, cnt, out, ker, ns); })
      .def("launch_sample_velocity",
           [](R&, Variable<1, float3> const& sx, Variable<1, float3> const& px,
              Variable<1, float> const& pd, Variable<1, float3> const& pv,
              Variable<2, float4> const& nbr, Variable<1, U> const& cnt, Variable<1, float3>& sv,
              Variable<2, float4> const& rec, Variable<2, float4> const& ker,
              Variable<1, float3> const& rx, Variable<1, float3> const& rv,
              Variable<1, float3> const& rw, U ns) {
             R::launch_sample_velocity(sx, px, pd, pv, nbr, cnt, sv, rec, ker, rx, rv, rw, ns);
           })
      .def("launch_sample_vorticity",
           [](R&, Variable<1, float3> const& sx, Variable<1, float3> const& px,
              Variable<1, float> const& pd, Variable<1, float3> const& pv,
              Variable<2, float4> const& nbr, Variable<1, U> const& cnt,
              Variable<1, float3> const& sv, Variable<1, float3>& vort,
              Variable<2, float4> const& rec, Variable<2, float4> const& ker,
              Variable<1, float3> const& rx, Variable<1, float3> const& rv,
              Variable<1, float3> const& rw, U ns) {
             R::launch_sample_vorticity(sx, px, pd, pv, nbr, cnt, sv, vort, rec, ker, rx, rv, rw, ns);
           })
      // the reference's partial-histogram and quantized buffers are accepted and unused
      .def("launch_histogram256",
           [](R&, py::object partial, Variable<1, U>& hist, py::object quantized4s,
              Variable<1, float> const& v, float v_min, float v_max, U n) {
             (void)partial;
             (void)quantized4s;
             R::launch_histogram256(v, v_min, v_max, n, hist);
           })
      .def("launch_histogram256_with_mask",
           [](R&, py::object partial, Variable<1, U>& hist, py::object quantized4s,
              Variable<1, float> const& mask, Variable<1, float> const& v, float v_min,
              float v_max, U n) {
             (void)partial;
             (void)quantized4s;
             R::launch_histogram256(v, v_min, v_max, n, hist, &mask);
           })
      .def_static("calculate_kl_divergence", &R::calculate_kl_divergence, py::arg("histogram_p"),
                  py::arg("histogram_q"), py::arg("n_p"), py::arg("n_q"),
                  py::arg("q_lower_bound"), py::arg("num_bins"))
      .def_static("sum", [](Variable<1, float> const& v, U n, U off) { return Runner<float>::sum(v, n, off); },
                  py::arg("var"), py::arg("n"), py::arg("offset") = 0)
      .def_static("sum", [](Variable<1, float3> const& v, U n, U off) { return Runner<float>::sum(v, n, off); },
                  py::arg("var"), py::arg("n"), py::arg("offset") = 0)
      .def_static("max", [](Variable<1, float> const& v, U n, U off) { return Runner<float>::max(v, n, off); },
                  py::arg("var"), py::arg("n"), py::arg("offset") = 0);

  // Distances are created with the reference policy: Pile::add takes
  // ownership (pile.hpp:113,303; pyalluvion.cu:612-614).
  py::module dg = m.def_submodule("dg", "analytic distances");
  using Dist = dg::Distance<float3, float>;
  py::class_<Dist>(dg, "Distancefloat")
      .def_property_readonly("aabb_min", &Dist::get_aabb_min)
      .def_property_readonly("aabb_max", &Dist::get_aabb_max);
  py::class_<dg::BoxDistance<float3, float>, Dist>(dg, "BoxDistancefloat")
      .def_static("create", &dg::BoxDistance<float3, float>::create, py::arg("widths"),
                  py::arg("outset") = 0, py::return_value_policy::reference);
  py::class_<dg::BoxShellDistance<float3, float>, Dist>(dg, "BoxShellDistancefloat")
      .def_static("create", &dg::BoxShellDistance<float3, float>::create,
                  py::arg("inner_widths"), py::arg("thickness"), py::arg("outset") = 0,
                  py::return_value_policy::reference);
  py::class_<dg::SphereDistance<float3, float>, Dist>(dg, "SphereDistancefloat")
      .def_static("create", &dg::SphereDistance<float3, float>::create, py::arg("radius"),
                  py::return_value_policy::reference);
  // A distance sampled on the map nodes (the reference's grid path, which it
  // takes for MeshDistance): signed_distance maps an [n, 3] float32 array of
  // body-frame points to n distances.
  py::class_<dg::MapDistance<float3, float>, Dist>(dg, "MapDistancefloat")
      .def_static(
          "create",
          [](float3 aabb_min, float3 aabb_max, py::function signed_distance) {
            auto fn = [signed_distance](const float* pts, unsigned n, float* out) {
              py::gil_scoped_acquire gil;
              py::array_t<float> arr({py::ssize_t(n), py::ssize_t(3)});
              std::memcpy(arr.mutable_data(), pts, size_t(n) * 3 * sizeof(float));
              py::array_t<float, py::array::c_style | py::array::forcecast> d =
                  signed_distance(arr);
              if (d.size() != py::ssize_t(n))
                throw std::invalid_argument("signed_distance must return one value per point");
              std::memcpy(out, d.data(), size_t(n) * sizeof(float));
            };
            return dg::MapDistance<float3, float>::create(aabb_min, aabb_max, fn);
          },
          py::arg("aabb_min"), py::arg("aabb_max"), py::arg("signed_distance"),
          py::return_value_policy::reference);

  py::class_<Mesh>(m, "Mesh")
      .def(py::init<>())
      .def_readonly("vertices", &Mesh::vertices)
      .def("set_vertices",
           [](Mesh& mesh, py::array_t<float, py::array::c_style | py::array::forcecast> xyz) {
             if (xyz.ndim() != 2 || xyz.shape(1) != 3)
               throw std::invalid_argument("vertices must be [n, 3]");
             mesh.vertices.resize(size_t(xyz.shape(0)));
             std::memcpy(mesh.vertices.data(), xyz.data(), mesh.vertices.size() * sizeof(float3));
           })
      .def("set_box", &Mesh::set_box, py::arg("widths"), py::arg("n") = 2)
      .def("set_uv_sphere", &Mesh::set_uv_sphere, py::arg("radius"), py::arg("num_sectors"),
           py::arg("num_stacks"))
      .def("translate", &Mesh::translate)
      .def("rotate", &Mesh::rotate)
      .def("scale", &Mesh::scale)
      .def("clear", &Mesh::clear);

  using P = Pile<float>;
  // Pile.x / v / q / omega: the reference binds its pinned kinematics
  // variables (pile.hpp:103-110, pyalluvion.cu:467-470), so pile.x[i] reads
  // and pile.x[i] = f3 writes body i.  This view does the same on the
  // library's host kinematics.
  struct PileKinematics {
    P* pile;
    int which;  // 0 x, 1 v, 2 q, 3 omega
    U size() const { return pile->get_size(); }
    U index(long i) const {
      long n = long(size());
      if (i < 0) i += n;
      if (i < 0 || i >= n) throw py::index_error("body index out of range");
      return U(i);
    }
    py::object get(long i) const {
      U b = index(i);
      switch (which) {
        case 0: return py::cast(pile->get_x(b));
        case 1: return py::cast(pile->get_v(b));
        case 2: return py::cast(pile->get_q(b));
        default: return py::cast(pile->get_omega(b));
      }
    }
    void set(long i, py::object value) const {
      U b = index(i);
      switch (which) {
        case 0: pile->set_x(b, value.cast<float3>()); break;
        case 1: pile->set_v(b, value.cast<float3>()); break;
        case 2: pile->set_q(b, value.cast<float4>()); break;
        default: pile->set_omega(b, value.cast<float3>()); break;
      }
    }
  };
  py::class_<PileKinematics>(m, "PileKinematics")
      .def("__len__", &PileKinematics::size)
      .def("__getitem__", &PileKinematics::get)
      .def("__setitem__", &PileKinematics::set)
      .def("__iter__",
           [](PileKinematics const& k) {
             py::list out;
             for (U i = 0; i < k.size(); ++i) out.append(k.get(long(i)));
             return py::iter(out);
           })
      .def("__repr__", [](PileKinematics const& k) {
        py::list out;
        for (U i = 0; i < k.size(); ++i) out.append(k.get(long(i)));
        return py::repr(out);
      });
  auto kin_property = [](int which) {
    return py::cpp_function([which](P& p) { return PileKinematics{&p, which}; },
                            py::keep_alive<0, 1>());
  };
  auto kin_setter = [](int which) {
    return py::cpp_function([which](P& p, py::sequence values) {
      PileKinematics k{&p, which};
      if (py::len(values) != k.size())
        throw std::invalid_argument("one value per body expected");
      for (U i = 0; i < k.size(); ++i) k.set(long(i), values[i]);
    });
  };
  // mass / restitution / friction / inertia_tensor lists (the reference's
  // std::vector members, pile.hpp:97-101): read as lists, assigned as lists
  auto body_getter = [](int which) {
    return py::cpp_function([which](P const& p) {
      py::list out;
      for (U i = 0; i < p.get_size(); ++i) {
        const auto b = p.get_body(i);
        if (which == 0) out.append(b.mass);
        if (which == 1) out.append(b.restitution);
        if (which == 2) out.append(b.friction);
        if (which == 3) out.append(b.inertia_tensor);
      }
      return out;
    });
  };
  auto body_setter = [](int which) {
    return py::cpp_function([which](P& p, py::sequence values) {
      if (py::len(values) != p.get_size())
        throw std::invalid_argument("one value per body expected");
      for (U i = 0; i < p.get_size(); ++i) {
        auto b = p.get_body(i);
        if (which == 0) b.mass = values[i].cast<float>();
        if (which == 1) b.restitution = values[i].cast<float>();
        if (which == 2) b.friction = values[i].cast<float>();
        if (which == 3) b.inertia_tensor = values[i].cast<float3>();
        p.set_body(i, b);
      }
    });
  };
  py::class_<P>(m, "Pilefloat")
      .def(py::init<Store&, Runner<float>&, U, VolumeMethod, U, Const<float>*, ConstiN*>(),
           py::arg("store"), py::arg("runner"), py::arg("max_num_contacts"),
           py::arg("volume_method") = VolumeMethod::volume_map,
           py::arg("max_num_pellets") = 10000, py::arg("cn") = nullptr, py::arg("cni") = nullptr,
           // the pile reads cn and cni later (contact tolerance, boundary count)
           py::keep_alive<1, 2>(), py::keep_alive<1, 7>(), py::keep_alive<1, 8>())
      .def("add",
           [](P& p, Dist* distance, U3 resolution, float cell_width, float margin, float sign,
              py::object collision_mesh, float mass, float restitution, float friction,
              float3 inertia_tensor, float3 x, float4 q, py::object display_mesh) {
             (void)display_mesh;
             Mesh none;
             const Mesh& coll = collision_mesh.is_none() ? none : collision_mesh.cast<const Mesh&>();
             return p.add(distance, resolution, cell_width, margin, sign, coll, mass, restitution,
                          friction, inertia_tensor, x, q);
           },
           py::arg("distance"), py::arg("resolution") = U3{0, 0, 0}, py::arg("cell_width") = 0,
           py::arg("margin") = std::numeric_limits<float>::infinity(), py::arg("sign") = 1,
           py::arg("collision_mesh") = py::none(), py::arg("mass") = 0,
           py::arg("restitution") = 1, py::arg("friction") = 0,
           py::arg("inertia_tensor") = float3{1, 1, 1}, py::arg("x") = float3{0, 0, 0},
           py::arg("q") = float4{0, 0, 0, 1}, py::arg("display_mesh") = py::none(),
           py::keep_alive<1, 2>())
      .def("add_pellets",
           [](P& p, Dist* distance, U3 resolution, float cell_width, float margin, float sign,
              py::array_t<float, py::array::c_style | py::array::forcecast> pellets, float mass,
              float restitution, float friction, float3 inertia_tensor, float3 x, float4 q,
              py::object display_mesh) {
             (void)display_mesh;
             if (pellets.ndim() != 2 || pellets.shape(1) != 3)
               throw std::invalid_argument("pellets must be [n, 3]");
             std::vector<float3> pts(size_t(pellets.shape(0)));
             std::memcpy(pts.data(), pellets.data(), pts.size() * sizeof(float3));
             return p.add_pellets(distance, resolution, cell_width, margin, sign, pts, mass,
                                  restitution, friction, inertia_tensor, x, q);
           },
           py::arg("distance"), py::arg("resolution") = U3{0, 0, 0}, py::arg("cell_width") = 0,
           py::arg("margin") = std::numeric_limits<float>::infinity(), py::arg("sign") = 1,
           py::arg("pellets"), py::arg("mass") = 0, py::arg("restitution") = 1,
           py::arg("friction") = 0, py::arg("inertia_tensor") = float3{1, 1, 1},
           py::arg("x") = float3{0, 0, 0}, py::arg("q") = float4{0, 0, 0, 1},
           py::arg("display_mesh") = py::none(), py::keep_alive<1, 2>())
      .def_property_readonly("num_pellets", &P::num_pellets)
      .def_property_readonly("num_collision_pellets", &P::num_pellets)
      .def_property_readonly("max_num_pellets", &P::max_num_pellets)
      .def("hint_identical_sequence", &P::hint_identical_sequence, py::arg("begin_id"),
           py::arg("end_id"))
      .def("get_size", &P::get_size)
      .def_property("mass", body_getter(0), body_setter(0))
      .def_property("restitution", body_getter(1), body_setter(1))
      .def_property("friction", body_getter(2), body_setter(2))
      .def_property("inertia_tensor", body_getter(3), body_setter(3))
      .def_property("x", kin_property(0), kin_setter(0))
      .def_property("v", kin_property(1), kin_setter(1))
      .def_property("q", kin_property(2), kin_setter(2))
      .def_property("omega", kin_property(3), kin_setter(3))
      .def("replace",
           [](P& p, U i, Dist* distance, U3 resolution, float cell_width, float margin,
              float sign, py::object collision_mesh, float mass, float restitution,
              float friction, float3 inertia_tensor, float3 x, float4 q,
              py::object display_mesh) {
             (void)display_mesh;
             Mesh none;
             const Mesh& coll = collision_mesh.is_none() ? none : collision_mesh.cast<const Mesh&>();
             p.replace(i, distance, resolution, cell_width, margin, sign, coll, mass, restitution,
                       friction, inertia_tensor, x, q);
           },
           py::arg("i"), py::arg("distance"), py::arg("resolution") = U3{0, 0, 0},
           py::arg("cell_width") = 0,
           py::arg("margin") = std::numeric_limits<float>::infinity(), py::arg("sign") = 1,
           py::arg("collision_mesh") = py::none(), py::arg("mass") = 0,
           py::arg("restitution") = 1, py::arg("friction") = 0,
           py::arg("inertia_tensor") = float3{1, 1, 1}, py::arg("x") = float3{0, 0, 0},
           py::arg("q") = float4{0, 0, 0, 1}, py::arg("display_mesh") = py::none(),
           py::keep_alive<1, 3>())
      .def("calculate_cfl_v2", &P::calculate_cfl_v2)
      .def("compute_mask", &P::compute_mask, py::arg("i"), py::arg("distance_threshold"),
           py::arg("sample_x"), py::arg("mask"), py::arg("num_samples"))
      .def("get_matrix",
           [](P const& p, U i) {
             // glm::mat4 is column-major: element (row r, column c) is m[4c + r]
             const auto m = p.get_matrix(i);
             py::array_t<float> out({4, 4});
             auto a = out.mutable_unchecked<2>();
             for (int r = 0; r < 4; ++r)
               for (int c = 0; c < 4; ++c) a(r, c) = m[4 * c + r];
             return out;
           },
           py::arg("i"))
      .def("get_x", &P::get_x)
      .def("get_v", &P::get_v)
      .def("get_q", &P::get_q)
      .def("get_omega", &P::get_omega)
      .def("get_force", &P::get_force)
      .def("get_torque", &P::get_torque)
      .def("set_x", &P::set_x)
      .def("set_v", &P::set_v)
      .def("set_q", &P::set_q)
      .def("set_omega", &P::set_omega)
      .def("set_gravity", &P::set_gravity)
      .def("integrate_kinematics", &P::integrate_kinematics)
      .def("copy_kinematics_to_device", &P::copy_kinematics_to_device)
      .def_property_readonly("x_device", py::cpp_function(&P::x_device, py::keep_alive<0, 1>()))
      .def_property_readonly("v_device", py::cpp_function(&P::v_device, py::keep_alive<0, 1>()))
      .def_property_readonly("q_device", py::cpp_function(&P::q_device, py::keep_alive<0, 1>()))
      .def_property_readonly("omega_device",
                             py::cpp_function(&P::omega_device, py::keep_alive<0, 1>()))
      .def("find_contacts", [](P& p, U i, U j) { return p.find_contacts(i, j); })
      .def("find_contacts", [](P& p, U i) { return p.find_contacts(i); })
      .def("find_contacts",
           [](P& p, const Variable<1, float3>& particle_x, U max_num_beads) {
             return p.find_contacts(particle_x, max_num_beads);
           },
           py::arg("particle_x"), py::arg("max_num_beads"))
      .def("find_contacts", [](P& p) { return p.find_contacts(); })
      .def("solve_contacts", &P::solve_contacts)
      .def_property_readonly("num_contacts_pinned",
                             [](P const& p) { return U(p.get_contacts().size()); })
      .def("get_contacts",
           [](P const& p) {
             py::list out;
             auto v3 = [](const float* a) { return py::make_tuple(a[0], a[1], a[2]); };
             for (const Contact& c : p.get_contacts()) {
               py::dict d;
               d["i"] = c.i;
               d["j"] = c.j;
               d["cp_i"] = v3(c.cp_i);
               d["cp_j"] = v3(c.cp_j);
               d["n"] = v3(c.n);
               d["t"] = v3(c.t);
               d["iiwi_diag"] = v3(c.iiwi_diag);
               d["iiwi_off_diag"] = v3(c.iiwi_off_diag);
               d["iiwj_diag"] = v3(c.iiwj_diag);
               d["iiwj_off_diag"] = v3(c.iiwj_off_diag);
               d["friction"] = c.friction;
               d["nkninv"] = c.nkninv;
               d["pmax"] = c.pmax;
               d["goalu"] = c.goalu;
               d["impulse_sum"] = c.impulse_sum;
               out.append(d);
             }
             return out;
           })
      .def("set_collision_vertices",
           [](P& p, U i, py::array_t<float, py::array::c_style | py::array::forcecast> xyz) {
             if (xyz.ndim() != 2 || xyz.shape(1) != 3)
               throw std::invalid_argument("vertices must be [n, 3]");
             check(alv_pile_set_collision_vertices(p.handle(), i, xyz.data(), U(xyz.shape(0))));
           })
      .def("write_file", &P::write_file, py::arg("filename"), py::arg("x_scale") = 1,
           py::arg("v_scale") = 1, py::arg("omega_scale") = 1)
      .def("read_file", &P::read_file, py::arg("filename"), py::arg("num_rigids") = -1,
           py::arg("dst_offset") = 0, py::arg("src_offset") = 0)
      .def_static("get_size_from_file", &P::get_size_from_file, py::arg("filename"))
      .def_property_readonly("volume_grids",
                             [](P const& p) {
                               std::vector<Variable<1, float>> out;
                               for (U i = 0; i < p.get_size(); ++i) out.push_back(p.get_volume_grid(i));
                               return out;
                             })
      .def_property_readonly("distance_grids", &P::get_distance_grids)
      .def_property_readonly("domain_min_list",
                             [](P const& p) {
                               std::vector<float3> out;
                               for (U i = 0; i < p.get_size(); ++i) out.push_back(p.get_grid_info(i).domain_min);
                               return out;
                             })
      .def_property_readonly("domain_max_list",
                             [](P const& p) {
                               std::vector<float3> out;
                               for (U i = 0; i < p.get_size(); ++i) out.push_back(p.get_grid_info(i).domain_max);
                               return out;
                             })
      .def_property_readonly("resolution_list",
                             [](P const& p) {
                               std::vector<U3> out;
                               for (U i = 0; i < p.get_size(); ++i) out.push_back(p.get_grid_info(i).resolution);
                               return out;
                             })
      .def_property_readonly("cell_size_list",
                             [](P const& p) {
                               std::vector<float3> out;
                               for (U i = 0; i < p.get_size(); ++i) out.push_back(p.get_grid_info(i).cell_size);
                               return out;
                             })
      .def_property_readonly("sign_list",
                             [](P const& p) {
                               std::vector<float> out;
                               for (U i = 0; i < p.get_size(); ++i) out.push_back(p.get_grid_info(i).sign);
                               return out;
                             })
      .def_property_readonly("grid_size_list", [](P const& p) {
        std::vector<U> out;
        for (U i = 0; i < p.get_size(); ++i) out.push_back(p.get_volume_grid(i).get_linear_shape());
        return out;
      });

  // Solver classes
  auto solver_common = [](auto& cls) {
    using S = typename std::remove_reference<decltype(cls)>::type::type;
    cls.def_readonly("max_num_particles", &S::max_num_particles)
        .def_property("num_particles", &S::get_num_particles, &S::set_num_particles)
        .def_property("t", &S::get_t, &S::set_t)
        .def_property("dt", &S::get_dt, &S::set_dt)
        .def_property("initial_dt", &S::get_initial_dt, &S::set_initial_dt)
        .def_property("max_dt", &S::get_max_dt, &S::set_max_dt)
        .def_property("min_dt", &S::get_min_dt, &S::set_min_dt)
        .def_property("cfl", &S::get_cfl, &S::set_cfl)
        .def_readonly("particle_radius", &S::particle_radius)
        .def_property_readonly("particle_max_v2", [](S const& s) { return s.report().particle_max_v2; })
        .def_property_readonly("pile_max_v2", [](S const& s) { return s.report().pile_max_v2; })
        .def_property_readonly("max_v2", [](S const& s) { return s.report().max_v2; })
        .def_property_readonly("cfl_dt", [](S const& s) { return s.report().cfl_dt; })
        .def_property_readonly("utilized_cfl", [](S const& s) { return s.report().utilized_cfl; })
        .def_property_readonly("particle_x", py::cpp_function(&S::particle_x, py::keep_alive<0, 1>()))
        .def_property_readonly("particle_v", py::cpp_function(&S::particle_v, py::keep_alive<0, 1>()))
        .def_property_readonly("particle_a", py::cpp_function(&S::particle_a, py::keep_alive<0, 1>()))
        .def_property_readonly("particle_density", py::cpp_function(&S::particle_density, py::keep_alive<0, 1>()))
        .def_property_readonly("particle_num_neighbors", py::cpp_function(&S::particle_num_neighbors, py::keep_alive<0, 1>()))
        .def_property_readonly("grid_anomaly", py::cpp_function(&S::grid_anomaly, py::keep_alive<0, 1>()))
        .def_property_readonly("boundary_mask", py::cpp_function(&S::boundary_mask, py::keep_alive<0, 1>()))
        .def_property_readonly("particle_neighbors", py::cpp_function(&S::particle_neighbors, py::keep_alive<0, 1>()))
        .def_property_readonly("particle_boundary", py::cpp_function(&S::particle_boundary, py::keep_alive<0, 1>()))
        .def_property_readonly("particle_boundary_kernel", py::cpp_function(&S::particle_boundary_kernel, py::keep_alive<0, 1>()))
        .def_property_readonly("particle_force", py::cpp_function(&S::particle_force, py::keep_alive<0, 1>()))
        .def_property_readonly("particle_torque", py::cpp_function(&S::particle_torque, py::keep_alive<0, 1>()))
        .def("set_static_boundary_forces", &S::set_static_boundary_forces, py::arg("enable"))
        .def_property_readonly("pid_length", py::cpp_function(&S::pid_length, py::keep_alive<0, 1>()))
        .def_property_readonly("pid", py::cpp_function(&S::pid, py::keep_alive<0, 1>()))
        .def("sample_all_boundaries", &S::sample_all_boundaries, py::arg("sample_x"),
             py::arg("sample_boundary"), py::arg("sample_boundary_kernel"), py::arg("num_samples"))
        .def("reset_solving_var", &S::reset_solving_var)
        .def("reset_t", &S::reset_t)
        .def("compute_all_boundaries", &S::compute_all_boundaries)
        .def("update_particle_neighbors", &S::template update_particle_neighbors<0>)
        .def("update_particle_neighbors_wrap1", &S::template update_particle_neighbors<1>)
        .def("compute_density_change", &S::compute_density_change)
        .def_property_readonly("num_scenes", &S::num_scenes)
        .def("set_graph", &S::set_graph, py::arg("enable"), py::arg("slack") = 0)
        .def("set_tiles", &S::set_tiles, py::arg("enable"))
        .def("scene_range", &S::scene_range, py::arg("scene"))
        .def("scene_report",
             [](S const& s, U k) {
               const alv_solver_report r = s.scene_report(k);
               py::dict d;
               d["num_divergence_solve"] = r.num_divergence_solve;
               d["num_density_solve"] = r.num_density_solve;
               d["mean_density_change"] = r.mean_density_change;
               d["mean_density_error"] = r.mean_density_error;
               d["max_v2"] = r.max_v2;
               d["cfl_dt"] = r.cfl_dt;
               d["utilized_cfl"] = r.utilized_cfl;
               d["t"] = r.t;
               d["dt"] = r.dt;
               return d;
             },
             py::arg("scene"))
        .def("set_warm_start_net", &S::set_warm_start_net, py::arg("net").none(true),
             py::arg("divergence") = true, py::arg("density") = true, py::keep_alive<1, 2>())
        .def("set_feature_log", &S::set_feature_log, py::arg("enable"))
        .def("export_training_pairs",
             [](S& s, py::object x, py::object y, size_t offset) {
               const size_t n = s.get_num_particles();
               float* px = var_floats(x, (offset + n) * 8, "x");
               float* py_ = var_floats(y, (offset + n) * 2, "y");
               s.export_training_pairs(px + offset * 8, py_ + offset * 2);
             },
             py::arg("x"), py::arg("y"), py::arg("offset") = 0)
        .def("integrate_velocity", &S::integrate_velocity)
        .def("step", &S::template step<0>)
        .def("step_wrap1", &S::template step<1>)
        .def("get_neighbors",
             [](S& s) {
               std::vector<U> ids, counts;
               s.get_neighbors(ids, counts);
               U K = U(ids.size() / std::max<size_t>(1, counts.size()));
               py::array_t<U> a({size_t(counts.size()), size_t(K)});
               std::memcpy(a.mutable_data(), ids.data(), ids.size() * sizeof(U));
               py::array_t<U> c(counts.size());
               std::memcpy(c.mutable_data(), counts.data(), counts.size() * sizeof(U));
               return py::make_tuple(a, c);
             })
        .def("transform_all_pellets", &S::transform_all_pellets)
        .def("get_pellet_neighbors",
             [](S& s, int which, U rows) {
               std::vector<U> ids, counts;
               s.get_pellet_neighbors(which, rows, ids, counts);
               U K = U(ids.size() / std::max<size_t>(1, counts.size()));
               py::array_t<U> a({size_t(counts.size()), size_t(K)});
               std::memcpy(a.mutable_data(), ids.data(), ids.size() * sizeof(U));
               py::array_t<U> c(counts.size());
               std::memcpy(c.mutable_data(), counts.data(), counts.size() * sizeof(U));
               return py::make_tuple(a, c);
             },
             py::arg("which"), py::arg("rows"))
        .def("get_grid", [](S& s) {
          std::vector<U> cnt, st, an;
          s.get_grid(cnt, st, an);
          return py::make_tuple(py::array_t<U>(cnt.size(), cnt.data()),
                                py::array_t<U>(st.size(), st.data()),
                                py::array_t<U>(an.size(), an.data()));
        });
  };

  // Solver base class (solver.hpp:8-306) carries the common fields.
  py::class_<SolverBase<float>> base(m, "Solverfloat");
  solver_common(base);

  using DF = SolverDf<float>;
  py::class_<DF, SolverBase<float>> df(m, "SolverDffloat");
  df.def(py::init<Runner<float>&, Pile<float>&, Store&, U, U, bool, bool, Const<float> const*,
                  ConstiN const*, bool>(),
         py::arg("runner"), py::arg("pile"), py::arg("store"), py::arg("max_num_particles"),
         py::arg("num_ushers") = 0, py::arg("enable_surface_tension") = false,
         py::arg("enable_vorticity") = false, py::arg("cn") = nullptr, py::arg("cni") = nullptr,
         py::arg("graphical") = false, py::keep_alive<1, 3>())
      // batch of independent scenes (no counterpart in the reference)
      .def(py::init([](Store& store, std::vector<Pile<float>*> piles, std::vector<U> num_particles,
                       std::vector<uint3> grid_res, std::vector<int3> grid_offset,
                       Const<float> const* cn, ConstiN const* cni) {
             const size_t n = piles.size();
             if (num_particles.size() != n || grid_res.size() != n || grid_offset.size() != n)
               throw std::invalid_argument("one entry per scene in every list");
             std::vector<SceneSpec<float>> sc(n);
             for (size_t k = 0; k < n; ++k)
               sc[k] = SceneSpec<float>{piles[k], num_particles[k], grid_res[k], grid_offset[k]};
             return new DF(store, sc, cn, cni);
           }),
           py::arg("store"), py::arg("piles"), py::arg("num_particles"), py::arg("grid_res"),
           py::arg("grid_offset"), py::arg("cn") = nullptr, py::arg("cni") = nullptr,
           py::keep_alive<1, 3>())
      .def_property_readonly("num_divergence_solve",
                             [](DF const& s) { return s.report().num_divergence_solve; })
      .def_property_readonly("num_density_solve", [](DF const& s) { return s.report().num_density_solve; })
      .def_property_readonly("mean_density_change",
                             [](DF const& s) { return s.report().mean_density_change; })
      .def_property_readonly("mean_density_error",
                             [](DF const& s) { return s.report().mean_density_error; })
      .def_property("enable_divergence_solve", &DF::get_enable_divergence_solve,
                    &DF::set_enable_divergence_solve)
      .def_property("enable_density_solve", &DF::get_enable_density_solve,
                    &DF::set_enable_density_solve)
      .def_property("density_change_tolerance", &DF::get_density_change_tolerance,
                    &DF::set_density_change_tolerance)
      .def_property("density_error_tolerance", &DF::get_density_error_tolerance,
                    &DF::set_density_error_tolerance)
      .def_property("min_density_solve", &DF::get_min_density_solve, &DF::set_min_density_solve)
      .def_property("max_density_solve", &DF::get_max_density_solve, &DF::set_max_density_solve)
      .def_property("min_divergence_solve", &DF::get_min_divergence_solve,
                    &DF::set_min_divergence_solve)
      .def_property("max_divergence_solve", &DF::get_max_divergence_solve,
                    &DF::set_max_divergence_solve)
      .def_property_readonly("particle_dfsph_factor", py::cpp_function(&DF::particle_dfsph_factor, py::keep_alive<0, 1>()))
      .def_property_readonly("particle_kappa", py::cpp_function(&DF::particle_kappa, py::keep_alive<0, 1>()))
      .def_property_readonly("particle_kappa_v", py::cpp_function(&DF::particle_kappa_v, py::keep_alive<0, 1>()))
      .def_property_readonly("particle_density_adv", py::cpp_function(&DF::particle_density_adv, py::keep_alive<0, 1>()));

  using II = SolverIi<float>;
  py::class_<II, SolverBase<float>> ii(m, "SolverIifloat");
  ii.def(py::init<Runner<float>&, Pile<float>&, Store&, U, U, bool, bool, Const<float> const*,
                  ConstiN const*, bool>(),
         py::arg("runner"), py::arg("pile"), py::arg("store"), py::arg("max_num_particles"),
         py::arg("num_ushers") = 0, py::arg("enable_surface_tension") = false,
         py::arg("enable_vorticity") = false, py::arg("cn") = nullptr, py::arg("cni") = nullptr,
         py::arg("graphical") = false, py::keep_alive<1, 3>())
      .def_property_readonly("num_density_solve", [](II const& s) { return s.report().num_density_solve; })
      .def_property_readonly("mean_density_error",
                             [](II const& s) { return s.report().mean_density_error; })
      .def_property("density_error_tolerance", &II::get_density_error_tolerance,
                    &II::set_density_error_tolerance)
      .def_property("min_density_solve", &II::get_min_density_solve, &II::set_min_density_solve)
      .def_property("max_density_solve", &II::get_max_density_solve, &II::set_max_density_solve)
      .def_property_readonly("particle_pressure", py::cpp_function(&II::particle_pressure, py::keep_alive<0, 1>()))
      .def_property_readonly("particle_last_pressure", py::cpp_function(&II::particle_last_pressure, py::keep_alive<0, 1>()))
      .def_property_readonly("particle_aii", py::cpp_function(&II::particle_aii, py::keep_alive<0, 1>()))
      .def_property_readonly("particle_dii", py::cpp_function(&II::particle_dii, py::keep_alive<0, 1>()))
      .def_property_readonly("particle_dij_pj", py::cpp_function(&II::particle_dij_pj, py::keep_alive<0, 1>()))
      .def_property_readonly("particle_adv_density", py::cpp_function(&II::particle_adv_density, py::keep_alive<0, 1>()))
      .def_property_readonly("particle_pressure_accel", py::cpp_function(&II::particle_pressure_accel, py::keep_alive<0, 1>()))
      .def_property_readonly("particle_density_err", py::cpp_function(&II::particle_density_err, py::keep_alive<0, 1>()));

  // SolverI: the settings, report and pressure field of SolverIi
  py::class_<SolverI<float>, II>(m, "SolverIfloat")
      .def(py::init<Runner<float>&, Pile<float>&, Store&, U, U, bool, bool, Const<float> const*,
                    ConstiN const*, bool>(),
           py::arg("runner"), py::arg("pile"), py::arg("store"), py::arg("max_num_particles"),
           py::arg("num_ushers") = 0, py::arg("enable_surface_tension") = false,
           py::arg("enable_vorticity") = false, py::arg("cn") = nullptr, py::arg("cni") = nullptr,
           py::arg("graphical") = false, py::keep_alive<1, 3>());

  py::class_<WarmStartNet>(m, "WarmStartNet")
      .def(py::init<unsigned long long>(), py::arg("seed") = 0)
      .def_static("num_params", &WarmStartNet::num_params)
      .def("get_params",
           [](WarmStartNet const& n) {
             auto p = n.get_params();
             return py::array_t<float>(p.size(), p.data());
           })
      .def("set_params",
           [](WarmStartNet& n, py::array_t<float, py::array::c_style | py::array::forcecast> p) {
             n.set_params(std::vector<float>(p.data(), p.data() + p.size()));
           })
      .def("get_normalization",
           [](WarmStartNet const& n) {
             auto v = n.get_normalization();
             return py::array_t<float>(v.size(), v.data());
           })
      .def("set_normalization",
           [](WarmStartNet& n, py::array_t<float, py::array::c_style | py::array::forcecast> v) {
             n.set_normalization(std::vector<float>(v.data(), v.data() + v.size()));
           })
      .def("fit_normalization",
           [](WarmStartNet& n, py::object x, py::object y, size_t count) {
             n.fit_normalization(var_floats(x, count * 8, "x"), var_floats(y, count * 2, "y"),
                                 count);
           })
      .def("forward",
           [](WarmStartNet& n, py::object x, py::object y, size_t count) {
             n.forward(var_floats(x, count * 8, "x"), count, var_floats(y, count * 2, "y"));
           })
      .def("loss_grad",
           [](WarmStartNet& n, py::object xn, py::object tn, size_t count) {
             std::vector<float> g;
             const double loss = n.loss_grad(var_floats(xn, count * 8, "xn"),
                                             var_floats(tn, count * 2, "tn"), count, g);
             return py::make_tuple(loss, py::array_t<float>(g.size(), g.data()));
           })
      .def("train",
           [](WarmStartNet& n, py::object x, py::object y, size_t count, U batch, U steps,
              float lr, unsigned long long seed) {
             auto l = n.train(var_floats(x, count * 8, "x"), var_floats(y, count * 2, "y"),
                              count, batch, steps, lr, seed);
             return py::array_t<float>(l.size(), l.data());
           },
           py::arg("x"), py::arg("y"), py::arg("count"), py::arg("batch") = 65536,
           py::arg("steps") = 1000, py::arg("lr") = 1e-3f, py::arg("seed") = 0);

  m.def("version", []() { return std::string(alv_version()); });
  // Per-kernel HIP-event timing (Runner::launch statistics, runner.hpp:4835-4859)
  m.def("profiling_enable", [](bool on) { check(alv_profiling_enable(on ? 1 : 0)); });
  m.def("profiling_reset", []() { check(alv_profiling_reset()); });
  m.def("profiling_stats", []() {
    unsigned n = 0;
    check(alv_profiling_num_kernels(&n));
    py::dict out;
    for (unsigned i = 0; i < n; ++i) {
      const char* name = nullptr;
      double ms = 0;
      unsigned long long launches = 0;
      check(alv_profiling_kernel(i, &name, &ms, &launches));
      out[py::str(name)] = py::make_tuple(ms, launches);
    }
    return out;
  });
  m.def("device_synchronize", []() { check(alv_device_synchronize()); });
}
