// bind.cpp -- pybind11 module novaphy_amd._core with the reference's Python
// names and signatures (python/bindings/bind_fluid.cpp:15-351, the World /
// ModelBuilder / SimState / Transform / PerformanceMonitor parts of
// bind_sim.cpp, bind_core.cpp, bind_math.cpp, bind_performance.cpp), plus
// numpy accessors and the C++-only PBFSettings fields.  Long GPU calls release
// the GIL.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "npx_host.hpp"

namespace py = pybind11;
using namespace novaphy;

// Vec3f <-> numpy float32 (3,) / any length-3 sequence.
namespace pybind11::detail {
template <>
struct type_caster<Vec3f> {
  PYBIND11_TYPE_CASTER(Vec3f, const_name("Vector3"));
  bool load(handle src, bool) {
    if (!src) return false;
    try {
      auto a = py::array_t<float, py::array::c_style | py::array::forcecast>::ensure(src);
      if (!a || a.size() != 3) return false;
      const float* p = a.data();
      value = Vec3f(p[0], p[1], p[2]);
      return true;
    } catch (...) {
      return false;
    }
  }
  static handle cast(const Vec3f& v, return_value_policy, handle) {
    py::array_t<float> a(3);
    auto m = a.mutable_unchecked<1>();
    for (int i = 0; i < 3; ++i) m(i) = v[i];
    return a.release();
  }
};
}  // namespace pybind11::detail

static py::array_t<float> quat_xyzw(const Quatf& q) {
  py::array_t<float> a(4);
  auto m = a.mutable_unchecked<1>();
  m(0) = q.x();
  m(1) = q.y();
  m(2) = q.z();
  m(3) = q.w();
  return a;
}
static Quatf quat_from(py::handle h) {
  auto a = py::array_t<float, py::array::c_style | py::array::forcecast>::ensure(h);
  if (!a || a.size() != 4) throw std::invalid_argument("quaternion must have 4 components [x, y, z, w]");
  const float* p = a.data();
  return Quatf(p[3], p[0], p[1], p[2]);
}
static py::array_t<float> vecs_numpy(const std::vector<Vec3f>& v) {
  py::array_t<float> a({(py::ssize_t)v.size(), (py::ssize_t)3});
  auto m = a.mutable_unchecked<2>();
  for (size_t i = 0; i < v.size(); ++i)
    for (int d = 0; d < 3; ++d) m(i, d) = v[i][d];
  return a;
}
static py::array_t<float> floats_numpy(const std::vector<float>& v) {
  py::array_t<float> a((py::ssize_t)v.size());
  std::copy(v.begin(), v.end(), a.mutable_data());
  return a;
}
static std::vector<Vec3f> vecs_from(py::handle h) {
  // accepts a (n, 3) array or a sequence of 3-vectors
  std::vector<Vec3f> out;
  if (py::isinstance<py::array>(h)) {
    auto a = py::array_t<float, py::array::c_style | py::array::forcecast>::ensure(h);
    if (a.ndim() == 2 && a.shape(1) == 3) {
      const float* p = a.data();
      out.resize(a.shape(0));
      for (py::ssize_t i = 0; i < a.shape(0); ++i) out[i] = Vec3f(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
      return out;
    }
  }
  for (auto item : h) out.push_back(item.cast<Vec3f>());
  return out;
}

PYBIND11_MODULE(_core, m) {
  m.doc() = "MI355X-native PBF fluid path with NovaPhy's FluidWorld / PBFSolver API";

  // ---------------------------------------------------------------- math
  py::class_<Transform>(m, "Transform")
      .def(py::init<>())
      .def(py::init([](const Vec3f& p, py::handle q) { return Transform(p, quat_from(q)); }), py::arg("position"),
           py::arg("rotation"))
      .def_readwrite("position", &Transform::position)
      .def_property(
          "rotation", [](const Transform& t) { return quat_xyzw(t.rotation); },
          [](Transform& t, py::handle q) { t.rotation = quat_from(q); })
      .def("transform_point", &Transform::transform_point, py::arg("point"))
      .def("transform_vector", &Transform::transform_vector, py::arg("vector"))
      .def("inverse", &Transform::inverse)
      .def("__mul__", &Transform::operator*, py::arg("other"))
      .def_static("identity", &Transform::identity)
      .def_static("from_translation", &Transform::from_translation, py::arg("t"))
      .def_static("from_rotation", [](py::handle q) { return Transform::from_rotation(quat_from(q)); }, py::arg("q"))
      .def_static("from_axis_angle", &Transform::from_axis_angle, py::arg("axis"), py::arg("angle"));

  // ---------------------------------------------------------------- core
  py::class_<RigidBody>(m, "RigidBody")
      .def(py::init<>())
      .def_readwrite("mass", &RigidBody::mass)
      .def_readwrite("com", &RigidBody::com)
      .def_property(
          "inertia",
          [](const RigidBody& b) {
            py::array_t<float> a({3, 3});
            auto mm = a.mutable_unchecked<2>();
            for (int i = 0; i < 3; ++i)
              for (int j = 0; j < 3; ++j) mm(i, j) = b.inertia(i, j);
            return a;
          },
          [](RigidBody& b, py::array_t<float, py::array::c_style | py::array::forcecast> a) {
            if (a.size() != 9) throw std::invalid_argument("inertia must be 3x3");
            for (int i = 0; i < 9; ++i) b.inertia(i / 3, i % 3) = a.data()[i];
          })
      .def("inv_mass", &RigidBody::inv_mass)
      .def("inv_inertia",
           [](const RigidBody& b) {
             const Mat3f I = b.inv_inertia();
             py::array_t<float> a({3, 3});
             auto mm = a.mutable_unchecked<2>();
             for (int i = 0; i < 3; ++i)
               for (int j = 0; j < 3; ++j) mm(i, j) = I(i, j);
             return a;
           })
      .def("is_static", &RigidBody::is_static)
      .def_static("from_box", &RigidBody::from_box, py::arg("mass"), py::arg("half_extents"))
      .def_static("from_sphere", &RigidBody::from_sphere, py::arg("mass"), py::arg("radius"))
      .def_static("make_static", &RigidBody::make_static);

  py::enum_<ShapeType>(m, "ShapeType")
      .value("Box", ShapeType::Box)
      .value("Sphere", ShapeType::Sphere)
      .value("Plane", ShapeType::Plane);

  py::class_<CollisionShape>(m, "CollisionShape")
      .def(py::init<>())
      .def_readwrite("type", &CollisionShape::type)
      .def_readwrite("body_index", &CollisionShape::body_index)
      .def_readwrite("local_transform", &CollisionShape::local_transform)
      .def_readwrite("friction", &CollisionShape::friction)
      .def_readwrite("restitution", &CollisionShape::restitution)
      .def_readwrite("half_extents", &CollisionShape::half_extents)
      .def_readwrite("radius", &CollisionShape::radius)
      .def_readwrite("normal", &CollisionShape::normal)
      .def_readwrite("offset", &CollisionShape::offset)
      // the reference's names for the primitive payloads (bind_core.cpp:182-205)
      .def_readwrite("box_half_extents", &CollisionShape::half_extents)
      .def_readwrite("sphere_radius", &CollisionShape::radius)
      .def_readwrite("plane_normal", &CollisionShape::normal)
      .def_readwrite("plane_offset", &CollisionShape::offset)
      .def_static("make_box", &CollisionShape::make_box, py::arg("half_extents"), py::arg("body_index"),
                  py::arg("local") = Transform::identity(), py::arg("friction") = 0.5f,
                  py::arg("restitution") = 0.3f)
      .def_static("make_sphere", &CollisionShape::make_sphere, py::arg("radius"), py::arg("body_index"),
                  py::arg("local") = Transform::identity(), py::arg("friction") = 0.5f,
                  py::arg("restitution") = 0.3f)
      .def_static("make_plane", &CollisionShape::make_plane, py::arg("normal"), py::arg("offset"),
                  py::arg("friction") = 0.5f, py::arg("restitution") = 0.0f);

  py::class_<Model>(m, "Model")
      .def_property_readonly("num_bodies", &Model::num_bodies)
      .def_property_readonly("num_shapes", &Model::num_shapes)
      .def_readonly("bodies", &Model::bodies)
      .def_readonly("shapes", &Model::shapes)
      .def_readonly("initial_transforms", &Model::initial_transforms);

  py::class_<ModelBuilder>(m, "ModelBuilder")
      .def(py::init<>())
      .def("add_body", &ModelBuilder::add_body, py::arg("body"), py::arg("transform") = Transform::identity())
      .def("add_shape", &ModelBuilder::add_shape, py::arg("shape"))
      .def("add_ground_plane", &ModelBuilder::add_ground_plane, py::arg("y") = 0.0f, py::arg("friction") = 0.5f,
           py::arg("restitution") = 0.0f)
      .def("build", &ModelBuilder::build)
      .def_property_readonly("num_bodies", &ModelBuilder::num_bodies)
      .def_property_readonly("num_shapes", &ModelBuilder::num_shapes);

  // ---------------------------------------------------------- collision
  // bind_core.cpp:94-120, 266-295; bind_collision.cpp:13-80
  py::class_<AABB>(m, "AABB")
      .def(py::init<>())
      .def(py::init<const Vec3f&, const Vec3f&>(), py::arg("min"), py::arg("max"))
      .def_readwrite("min", &AABB::min)
      .def_readwrite("max", &AABB::max)
      .def("overlaps", &AABB::overlaps, py::arg("other"))
      .def_static("from_sphere", &AABB::from_sphere, py::arg("center"), py::arg("radius"))
      .def_static("from_oriented_box", &AABB::from_oriented_box, py::arg("half_extents"), py::arg("transform"));
  py::class_<ContactPoint>(m, "ContactPoint")
      .def(py::init<>())
      .def_readwrite("position", &ContactPoint::position)
      .def_readwrite("normal", &ContactPoint::normal)
      .def_readwrite("penetration", &ContactPoint::penetration)
      .def_readwrite("body_a", &ContactPoint::body_a)
      .def_readwrite("body_b", &ContactPoint::body_b)
      .def_readwrite("shape_a", &ContactPoint::shape_a)
      .def_readwrite("shape_b", &ContactPoint::shape_b)
      .def_readwrite("friction", &ContactPoint::friction)
      .def_readwrite("restitution", &ContactPoint::restitution)
      .def_readwrite("accumulated_normal_impulse", &ContactPoint::accumulated_normal_impulse)
      .def_readwrite("accumulated_tangent_impulse_1", &ContactPoint::accumulated_tangent_impulse_1)
      .def_readwrite("accumulated_tangent_impulse_2", &ContactPoint::accumulated_tangent_impulse_2);
  py::class_<BroadPhasePair>(m, "BroadPhasePair")
      .def(py::init<>())
      .def_readwrite("body_a", &BroadPhasePair::body_a)
      .def_readwrite("body_b", &BroadPhasePair::body_b);
  py::class_<SweepAndPrune>(m, "SweepAndPrune")
      .def(py::init<>())
      .def("update", &SweepAndPrune::update, py::arg("body_aabbs"), py::arg("static_mask"))
      .def("get_pairs", &SweepAndPrune::get_pairs);
  m.def(
      "collide_shapes",
      [](const CollisionShape& a, const Transform& ta, const CollisionShape& b, const Transform& tb) {
        std::vector<ContactPoint> contacts;
        const bool hit = collide_shapes(a, ta, b, tb, contacts);
        return std::make_pair(hit, contacts);
      },
      py::arg("shape_a"), py::arg("transform_a"), py::arg("shape_b"), py::arg("transform_b"));

  // ---------------------------------------------------------------- sim
  py::class_<SolverSettings>(m, "SolverSettings")
      .def(py::init<>())
      .def_readwrite("velocity_iterations", &SolverSettings::velocity_iterations)
      .def_readwrite("baumgarte", &SolverSettings::baumgarte)
      .def_readwrite("slop", &SolverSettings::slop)
      .def_readwrite("warm_starting", &SolverSettings::warm_starting)
      .def_readwrite("sleep_enabled", &SolverSettings::sleep_enabled)
      .def_readwrite("sleep_energy_threshold", &SolverSettings::sleep_energy_threshold)
      .def_readwrite("sleep_time_required", &SolverSettings::sleep_time_required)
      .def_readwrite("sleep_ema_alpha", &SolverSettings::sleep_ema_alpha);

  py::class_<SimState>(m, "SimState")
      .def(py::init<>())
      .def_readonly("transforms", &SimState::transforms)
      .def_readonly("linear_velocities", &SimState::linear_velocities)
      .def_readonly("angular_velocities", &SimState::angular_velocities)
      .def("set_transform",
           [](SimState& s, int i, const Transform& t) {
             if (i < 0 || i >= (int)s.transforms.size()) throw py::index_error("body index out of range");
             s.transforms[i] = t;
           },
           py::arg("body_index"), py::arg("transform"))
      .def("set_linear_velocity", &SimState::set_linear_velocity, py::arg("body_index"), py::arg("velocity"))
      .def("set_angular_velocity", &SimState::set_angular_velocity, py::arg("body_index"), py::arg("velocity"))
      .def("apply_force", &SimState::apply_force, py::arg("body_index"), py::arg("force"))
      .def("apply_torque", &SimState::apply_torque, py::arg("body_index"), py::arg("torque"));

  py::class_<PerformancePhaseStat>(m, "PerformancePhaseStat")
      .def(py::init<>())
      .def_readonly("name", &PerformancePhaseStat::name)
      .def_readonly("last_ms", &PerformancePhaseStat::last_ms)
      .def_readonly("avg_ms", &PerformancePhaseStat::avg_ms)
      .def_readonly("max_ms", &PerformancePhaseStat::max_ms)
      .def_readonly("samples", &PerformancePhaseStat::samples);
  py::class_<PerformanceMetric>(m, "PerformanceMetric")
      .def(py::init<>())
      .def_readonly("name", &PerformanceMetric::name)
      .def_readonly("value", &PerformanceMetric::value);
  py::class_<PerformanceMonitor>(m, "PerformanceMonitor")
      .def(py::init<>())
      .def_readwrite("enabled", &PerformanceMonitor::enabled)
      .def_readwrite("trace_enabled", &PerformanceMonitor::trace_enabled)
      .def_property(
          "trace_frame_capacity", [](const PerformanceMonitor& p) { return p.trace_frame_capacity; },
          [](PerformanceMonitor& p, int c) { p.trace_frame_capacity = std::max(0, c); })
      .def_property_readonly("frame_count", &PerformanceMonitor::frame_count)
      .def_property_readonly("last_frame_total_ms", &PerformanceMonitor::last_frame_total_ms)
      .def("reset", &PerformanceMonitor::reset)
      .def("phase_stats", &PerformanceMonitor::phase_stats)
      .def("last_frame_metrics", &PerformanceMonitor::last_frame_metrics)
      .def("write_trace_json", [](const PerformanceMonitor& p, py::object path) {
        p.write_trace_json(py::str(path).cast<std::string>());
      }, py::arg("output_path"));

  py::class_<World>(m, "World")
      .def(py::init<const Model&, SolverSettings>(), py::arg("model"), py::arg("solver_settings") = SolverSettings{})
      .def("step", &World::step, py::arg("dt"), py::call_guard<py::gil_scoped_release>())
      .def("set_gravity", &World::set_gravity, py::arg("gravity"))
      .def_property_readonly("gravity", &World::gravity)
      .def_property_readonly("state", py::overload_cast<>(&World::state), py::return_value_policy::reference_internal)
      .def_property_readonly("model", &World::model, py::return_value_policy::reference_internal)
      .def_property_readonly("contacts", &World::contacts)
      .def_property_readonly("performance_monitor", &World::performance_monitor,
                             py::return_value_policy::reference_internal)
      .def_property_readonly("solver_settings", &World::solver_settings,
                             py::return_value_policy::reference_internal)
      .def("apply_force", &World::apply_force, py::arg("body_index"), py::arg("force"))
      .def("apply_torque", &World::apply_torque, py::arg("body_index"), py::arg("torque"));

  // A coupling round over many rigid worlds in one call each way, for a
  // server stepping many coupled scenes: the poses of every body of every
  // world (world by world, rows of position, w, x, y, z: the pose layout of
  // npx_batch_coupling), and the coupling forces handed to each world followed
  // by its rigid step, with FluidWorld::step's rules (static bodies, and
  // bodies whose force and torque both have a zero float32 squared norm, get
  // nothing).
  m.def(
      "gather_poses7",
      [](const std::vector<World*>& worlds) {
        size_t nb = 0;
        for (const World* w : worlds) nb += w->state().transforms.size();
        py::array_t<float> out({(py::ssize_t)nb, (py::ssize_t)7});
        float* p = out.mutable_data();
        for (const World* w : worlds)
          for (const Transform& t : w->state().transforms) {
            p[0] = t.position[0];
            p[1] = t.position[1];
            p[2] = t.position[2];
            p[3] = t.rotation.w();
            p[4] = t.rotation.x();
            p[5] = t.rotation.y();
            p[6] = t.rotation.z();
            p += 7;
          }
        return out;
      },
      py::arg("worlds"));
  m.def(
      "apply_forces_and_step",
      [](const std::vector<World*>& worlds, py::array_t<float, py::array::c_style | py::array::forcecast> out6,
         const std::vector<float>& dts) {
        if (dts.size() != worlds.size()) throw std::invalid_argument("one dt per world expected");
        size_t nb = 0;
        for (const World* w : worlds) nb += w->state().transforms.size();
        if (out6.ndim() != 2 || out6.shape(1) != 6 || (size_t)out6.shape(0) != nb)
          throw std::invalid_argument("forces must be (" + std::to_string(nb) + ", 6): force and torque per body");
        const float* f = out6.data();
        py::gil_scoped_release nogil;
        for (size_t w = 0; w < worlds.size(); ++w) {
          World& W = *worlds[w];
          const int n = (int)W.state().transforms.size();
          for (int k = 0; k < n; ++k, f += 6) {
            if (W.model().bodies[k].is_static()) continue;
            const Vec3f F(f[0], f[1], f[2]), T(f[3], f[4], f[5]);
            if (F.squaredNorm() == 0.0f && T.squaredNorm() == 0.0f) continue;
            W.apply_force(k, F);
            W.apply_torque(k, T);
          }
          W.step(dts[w]);
        }
      },
      py::arg("worlds"), py::arg("forces"), py::arg("dts"));

  // ---------------------------------------------------------------- fluid
  py::class_<SPHKernels>(m, "SPHKernels")
      .def_static("poly6", &SPHKernels::poly6, py::arg("r_sq"), py::arg("h"))
      .def_static("spiky_grad", &SPHKernels::spiky_grad, py::arg("r"), py::arg("h"))
      .def_static("spiky", &SPHKernels::spiky, py::arg("r_len"), py::arg("h"))
      .def_static("cubic_spline", &SPHKernels::cubic_spline, py::arg("r_len"), py::arg("h"));

  py::class_<FluidBlockDef>(m, "FluidBlockDef")
      .def(py::init<>())
      .def_readwrite("lower", &FluidBlockDef::lower)
      .def_readwrite("upper", &FluidBlockDef::upper)
      .def_readwrite("particle_spacing", &FluidBlockDef::particle_spacing)
      .def_readwrite("rest_density", &FluidBlockDef::rest_density)
      .def_readwrite("initial_velocity", &FluidBlockDef::initial_velocity);

  py::class_<ParticleState>(m, "ParticleState")
      .def(py::init<>())
      .def("init", [](ParticleState& s, py::handle pos, const Vec3f& v0) { s.init(vecs_from(pos), v0); },
           py::arg("initial_positions"), py::arg("initial_velocity") = Vec3f::Zero())
      .def("init_arrays",
           [](ParticleState& s, py::array_t<float, py::array::c_style | py::array::forcecast> pos,
              py::object vel) {
             if (pos.ndim() != 2 || pos.shape(1) != 3) throw std::invalid_argument("positions must be (n, 3)");
             if (vel.is_none()) {
               s.init_arrays(pos.data(), nullptr, (int)pos.shape(0));
             } else {
               auto v = py::array_t<float, py::array::c_style | py::array::forcecast>::ensure(vel);
               if (!v || v.size() != pos.size()) throw std::invalid_argument("velocities must match positions");
               s.init_arrays(pos.data(), v.data(), (int)pos.shape(0));
             }
           },
           py::arg("positions"), py::arg("velocities") = py::none())
      .def("clear", &ParticleState::clear)
      .def_property_readonly("num_particles", &ParticleState::num_particles)
      .def_property_readonly("positions", [](const ParticleState& s) { return s.positions(); })
      .def_property_readonly("velocities", [](const ParticleState& s) { return s.velocities(); })
      .def_property_readonly("densities", [](const ParticleState& s) { return s.densities(); })
      .def_property_readonly("lambdas", [](const ParticleState& s) { return s.lambdas(); })
      .def("positions_numpy", [](const ParticleState& s) { return vecs_numpy(s.positions()); })
      .def("velocities_numpy", [](const ParticleState& s) { return vecs_numpy(s.velocities()); })
      .def("densities_numpy", [](const ParticleState& s) { return floats_numpy(s.densities()); })
      .def("lambdas_numpy", [](const ParticleState& s) { return floats_numpy(s.lambdas()); });

  py::class_<SpatialHashGrid>(m, "SpatialHashGrid")
      .def(py::init<float>(), py::arg("cell_size") = 0.1f)
      .def("build", [](SpatialHashGrid& g, py::handle pos) { g.build(vecs_from(pos)); }, py::arg("positions"))
      .def("clear", &SpatialHashGrid::clear)
      .def("query_neighbors", &SpatialHashGrid::query_neighbors, py::arg("point"), py::arg("radius"))
      .def("query_neighbors_batch",
           [](const SpatialHashGrid& g, py::handle q) { return g.query_neighbors_batch(vecs_from(q)); },
           py::arg("points"))
      .def("query_all_pairs",
           [](const SpatialHashGrid& g, py::handle pos, float r) { return g.query_all_pairs(vecs_from(pos), r); },
           py::arg("positions"), py::arg("radius"))
      .def_property("cell_size", &SpatialHashGrid::cell_size, &SpatialHashGrid::set_cell_size);

  py::class_<PBFSettings>(m, "PBFSettings")
      .def(py::init<>())
      .def_readwrite("rest_density", &PBFSettings::rest_density)
      .def_readwrite("kernel_radius", &PBFSettings::kernel_radius)
      .def_readwrite("particle_radius", &PBFSettings::particle_radius)
      .def_readwrite("solver_iterations", &PBFSettings::solver_iterations)
      .def_readwrite("epsilon", &PBFSettings::epsilon)
      .def_readwrite("xsph_viscosity", &PBFSettings::xsph_viscosity)
      .def_readwrite("vorticity_epsilon", &PBFSettings::vorticity_epsilon)
      .def_readwrite("corr_delta_q", &PBFSettings::corr_delta_q)
      .def_readwrite("corr_k", &PBFSettings::corr_k)
      .def_readwrite("corr_n", &PBFSettings::corr_n)
      .def_readwrite("use_domain_bounds", &PBFSettings::use_domain_bounds)
      .def_readwrite("domain_lower", &PBFSettings::domain_lower)
      .def_readwrite("domain_upper", &PBFSettings::domain_upper)
      .def("particle_mass", &PBFSettings::particle_mass, py::arg("spacing"));

  py::class_<PBFSolver>(m, "PBFSolver")
      .def(py::init<const PBFSettings&, uint32_t>(), py::arg("settings") = PBFSettings{}, py::arg("flags") = 0u)
      .def("step", &PBFSolver::step, py::arg("state"), py::arg("dt"), py::arg("gravity"), py::arg("particle_mass"),
           py::call_guard<py::gil_scoped_release>())
      .def("step_n", &PBFSolver::step_n, py::arg("state"), py::arg("dt"), py::arg("gravity"),
           py::arg("particle_mass"), py::arg("nsteps"), py::call_guard<py::gil_scoped_release>())
      .def_property_readonly("settings", py::overload_cast<>(&PBFSolver::settings),
                             py::return_value_policy::reference_internal);

  py::class_<BoundaryParticle>(m, "BoundaryParticle")
      .def(py::init<>())
      .def_readwrite("local_position", &BoundaryParticle::local_position)
      .def_readwrite("body_index", &BoundaryParticle::body_index)
      .def_readwrite("volume", &BoundaryParticle::volume);

  py::class_<FluidWorld, World>(m, "FluidWorld")
      .def(py::init<const Model&, const std::vector<FluidBlockDef>&, SolverSettings, PBFSettings, float, uint32_t>(),
           py::arg("model"), py::arg("fluid_blocks") = std::vector<FluidBlockDef>{},
           py::arg("solver_settings") = SolverSettings{}, py::arg("pbf_settings") = PBFSettings{},
           py::arg("boundary_extent") = 1.0f, py::arg("flags") = 0u)
      .def("step", &FluidWorld::step, py::arg("dt"), py::call_guard<py::gil_scoped_release>())
      .def_property_readonly("fluid_state", &FluidWorld::fluid_state, py::return_value_policy::reference_internal)
      .def_property_readonly("num_particles", &FluidWorld::num_particles)
      .def_property_readonly("num_boundary_particles", &FluidWorld::num_boundary_particles)
      .def_property_readonly("boundary_particles", &FluidWorld::boundary_particles)
      .def_property_readonly("pbf_settings", &FluidWorld::pbf_settings, py::return_value_policy::reference_internal)
      .def_property_readonly("particle_mass", &FluidWorld::particle_mass);

  m.def("generate_fluid_block", &generate_fluid_block, py::arg("block_def"));
  m.def("sample_model_boundaries", &sample_model_boundaries, py::arg("model"), py::arg("spacing"),
        py::arg("plane_extent") = 1.0f);
  m.def("device_count", &npx_device_count);
  m.attr("FLAG_FAST_MATH") = (uint32_t)NPX_FLAG_FAST_MATH;
}
