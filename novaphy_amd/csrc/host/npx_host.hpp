// npx_host.hpp -- C++20 host of the MI355X PBF path.
//
// Keeps the reference's class API (FluidWorld, PBFSolver, PBFSettings,
// ParticleState, SpatialHashGrid, SPHKernels, BoundaryParticle, ModelBuilder,
// Model, RigidBody, CollisionShape, Transform, World, SimState,
// PerformanceMonitor) so the GPU path is a drop-in for the reference's CPU
// one.  Every fluid computation goes through libnpx_hip's C ABI
// (include/novaphy_hip.h); the rigid-body pipeline stays on the host.
#pragma once

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <span>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

#include "novaphy_hip.h"

namespace novaphy {

constexpr float PI = 3.14159265358979323846f;  // math_utils.h:53

// ------------------------------------------------------------------ math
// Small fixed vectors with Eigen's evaluation order where the reference's
// numerics depend on it (squaredNorm = x*x + (y*y + z*z)).
struct Vec3f {
  float v[3] = {0.0f, 0.0f, 0.0f};
  Vec3f() = default;
  Vec3f(float x, float y, float z) : v{x, y, z} {}
  float& operator[](int i) { return v[i]; }
  float operator[](int i) const { return v[i]; }
  float x() const { return v[0]; }
  float y() const { return v[1]; }
  float z() const { return v[2]; }
  static Vec3f Zero() { return {}; }
  Vec3f operator+(const Vec3f& o) const { return {v[0] + o.v[0], v[1] + o.v[1], v[2] + o.v[2]}; }
  Vec3f operator-(const Vec3f& o) const { return {v[0] - o.v[0], v[1] - o.v[1], v[2] - o.v[2]}; }
  Vec3f operator-() const { return {-v[0], -v[1], -v[2]}; }
  Vec3f operator*(float s) const { return {v[0] * s, v[1] * s, v[2] * s}; }
  Vec3f& operator+=(const Vec3f& o) {
    for (int i = 0; i < 3; ++i) v[i] += o.v[i];
    return *this;
  }
  Vec3f& operator-=(const Vec3f& o) {
    for (int i = 0; i < 3; ++i) v[i] -= o.v[i];
    return *this;
  }
  Vec3f& operator*=(float s) {
    for (int i = 0; i < 3; ++i) v[i] *= s;
    return *this;
  }
  float dot(const Vec3f& o) const { return v[0] * o.v[0] + (v[1] * o.v[1] + v[2] * o.v[2]); }
  float squaredNorm() const { return dot(*this); }
  float norm() const { return std::sqrt(squaredNorm()); }
  Vec3f cross(const Vec3f& b) const {
    return {v[1] * b.v[2] - v[2] * b.v[1], v[2] * b.v[0] - v[0] * b.v[2], v[0] * b.v[1] - v[1] * b.v[0]};
  }
  Vec3f normalized() const {
    float z = squaredNorm();
    if (z > 0.0f) {
      float l = std::sqrt(z);
      return {v[0] / l, v[1] / l, v[2] / l};
    }
    return *this;
  }
  Vec3f operator/(float s) const { return {v[0] / s, v[1] / s, v[2] / s}; }
  Vec3f cwiseMin(const Vec3f& o) const {
    return {std::min(v[0], o.v[0]), std::min(v[1], o.v[1]), std::min(v[2], o.v[2])};
  }
  Vec3f cwiseMax(const Vec3f& o) const {
    return {std::max(v[0], o.v[0]), std::max(v[1], o.v[1]), std::max(v[2], o.v[2])};
  }
  static Vec3f Constant(float c) { return {c, c, c}; }
};
inline Vec3f operator*(float s, const Vec3f& a) { return {s * a.v[0], s * a.v[1], s * a.v[2]}; }

struct Mat3f {
  float m[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  static Mat3f Zero() { return {}; }
  static Mat3f Identity() {
    Mat3f r;
    r.m[0][0] = r.m[1][1] = r.m[2][2] = 1.0f;
    return r;
  }
  float& operator()(int i, int j) { return m[i][j]; }
  float operator()(int i, int j) const { return m[i][j]; }
  // Products sum their three terms as a0 + (a1 + a2), the order of Eigen's
  // unrolled reduction for fixed size 3 (the same as dot / squaredNorm).
  Vec3f operator*(const Vec3f& a) const {
    Vec3f r;
    for (int i = 0; i < 3; ++i) r[i] = m[i][0] * a[0] + (m[i][1] * a[1] + m[i][2] * a[2]);
    return r;
  }
  Mat3f operator*(const Mat3f& b) const {
    Mat3f r;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) r.m[i][j] = m[i][0] * b.m[0][j] + (m[i][1] * b.m[1][j] + m[i][2] * b.m[2][j]);
    return r;
  }
  Mat3f operator*(float s) const {
    Mat3f r = *this;
    for (auto& row : r.m)
      for (auto& e : row) e *= s;
    return r;
  }
  Vec3f col(int j) const { return {m[0][j], m[1][j], m[2][j]}; }
  Mat3f transpose() const {
    Mat3f r;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) r.m[i][j] = m[j][i];
    return r;
  }
  Mat3f inverse() const;
};

// Unit quaternion (w, x, y, z); rotation of a vector in Eigen's form.
struct Quatf {
  float w_ = 1.0f, x_ = 0.0f, y_ = 0.0f, z_ = 0.0f;
  Quatf() = default;
  Quatf(float w, float x, float y, float z) : w_(w), x_(x), y_(y), z_(z) {}
  static Quatf Identity() { return {}; }
  float w() const { return w_; }
  float x() const { return x_; }
  float y() const { return y_; }
  float z() const { return z_; }
  Quatf normalized() const {
    float n = std::sqrt(w_ * w_ + (x_ * x_ + (y_ * y_ + z_ * z_)));
    return {w_ / n, x_ / n, y_ / n, z_ / n};
  }
  Quatf conjugate() const { return {w_, -x_, -y_, -z_}; }
  Quatf operator*(const Quatf& b) const {
    return {w_ * b.w_ - x_ * b.x_ - y_ * b.y_ - z_ * b.z_, w_ * b.x_ + x_ * b.w_ + y_ * b.z_ - z_ * b.y_,
            w_ * b.y_ + y_ * b.w_ + z_ * b.x_ - x_ * b.z_, w_ * b.z_ + z_ * b.w_ + x_ * b.y_ - y_ * b.x_};
  }
  Vec3f rotate(const Vec3f& v) const {
    Vec3f qv(x_, y_, z_);
    Vec3f uv = qv.cross(v);
    uv = uv + uv;
    return (v + w_ * uv) + qv.cross(uv);
  }
  Mat3f matrix() const;
  static Quatf from_axis_angle(const Vec3f& axis, float angle) {
    Vec3f a = axis.normalized();
    float s = std::sin(0.5f * angle);
    return {std::cos(0.5f * angle), a[0] * s, a[1] * s, a[2] * s};
  }
};

// Transform (math_types.h:38-135): translation + unit quaternion.
struct Transform {
  Vec3f position;
  Quatf rotation;
  Transform() = default;
  Transform(const Vec3f& p, const Quatf& r) : position(p), rotation(r.normalized()) {}
  Transform operator*(const Transform& o) const {
    return Transform(position + rotation.rotate(o.position), rotation * o.rotation);
  }
  Transform inverse() const {
    Quatf inv = rotation.conjugate();
    return Transform(inv.rotate(-position), inv);
  }
  Vec3f transform_point(const Vec3f& p) const { return position + rotation.rotate(p); }
  Vec3f transform_vector(const Vec3f& v) const { return rotation.rotate(v); }
  Mat3f rotation_matrix() const { return rotation.matrix(); }
  static Transform identity() { return {}; }
  static Transform from_translation(const Vec3f& t) { return Transform(t, Quatf::Identity()); }
  static Transform from_rotation(const Quatf& q) { return Transform(Vec3f::Zero(), q); }
  static Transform from_axis_angle(const Vec3f& axis, float angle) {
    return Transform(Vec3f::Zero(), Quatf::from_axis_angle(axis, angle));
  }
};

// ------------------------------------------------------------------ model
struct RigidBody {  // body.h:15-103
  float mass = 1.0f;
  Vec3f com;
  Mat3f inertia = Mat3f::Identity();
  float inv_mass() const { return mass > 0.0f ? 1.0f / mass : 0.0f; }
  Mat3f inv_inertia() const { return mass <= 0.0f ? Mat3f::Zero() : inertia.inverse(); }
  bool is_static() const { return mass <= 0.0f; }
  static RigidBody from_box(float m, const Vec3f& he);
  static RigidBody from_sphere(float m, float radius);
  static RigidBody make_static();
};

enum class ShapeType { Box, Sphere, Plane };

struct CollisionShape {  // shape.h:43-142
  ShapeType type = ShapeType::Sphere;
  int body_index = -1;
  Transform local_transform;
  float friction = 0.5f;
  float restitution = 0.3f;
  Vec3f half_extents{0.5f, 0.5f, 0.5f};  // box
  float radius = 0.5f;                   // sphere
  Vec3f normal{0.0f, 1.0f, 0.0f};        // plane
  float offset = 0.0f;                   // plane
  static CollisionShape make_box(const Vec3f& he, int body, const Transform& local = Transform::identity(),
                                 float friction = 0.5f, float restitution = 0.3f);
  static CollisionShape make_sphere(float r, int body, const Transform& local = Transform::identity(),
                                    float friction = 0.5f, float restitution = 0.3f);
  static CollisionShape make_plane(const Vec3f& n, float offset, float friction = 0.5f, float restitution = 0.0f);
};

struct Model {  // model.h:17-35
  std::vector<RigidBody> bodies;
  std::vector<Transform> initial_transforms;
  std::vector<CollisionShape> shapes;
  int num_bodies() const { return (int)bodies.size(); }
  int num_shapes() const { return (int)shapes.size(); }
};

class ModelBuilder {  // model_builder.h:33-58
 public:
  int add_body(const RigidBody& b, const Transform& t = Transform::identity()) {
    bodies_.push_back(b);
    transforms_.push_back(t);
    return (int)bodies_.size() - 1;
  }
  int add_shape(const CollisionShape& s) {
    shapes_.push_back(s);
    return (int)shapes_.size() - 1;
  }
  int add_ground_plane(float y = 0.0f, float friction = 0.5f, float restitution = 0.0f) {
    return add_shape(CollisionShape::make_plane(Vec3f(0, 1, 0), y, friction, restitution));
  }
  Model build() const { return Model{bodies_, transforms_, shapes_}; }
  int num_bodies() const { return (int)bodies_.size(); }
  int num_shapes() const { return (int)shapes_.size(); }

 private:
  std::vector<RigidBody> bodies_;
  std::vector<Transform> transforms_;
  std::vector<CollisionShape> shapes_;
};

// ------------------------------------------------------------ sim
struct SolverSettings {  // free_body_solver.h:20-31
  int velocity_iterations = 30;
  float baumgarte = 0.3f;
  float slop = 0.005f;
  bool warm_starting = true;
  bool sleep_enabled = false;
  float sleep_energy_threshold = 0.01f;
  float sleep_time_required = 0.5f;
  float sleep_ema_alpha = 0.8f;
};

struct ContactPoint {  // contact.h:17-40; normal points from body A toward body B
  Vec3f position, normal;
  float penetration = 0.0f;
  int body_a = -1, body_b = -1, shape_a = -1, shape_b = -1;
  float friction = 0.5f, restitution = 0.0f;
  float accumulated_normal_impulse = 0.0f;
  float accumulated_tangent_impulse_1 = 0.0f;
  float accumulated_tangent_impulse_2 = 0.0f;
};

inline float combine_friction(float a, float b) { return std::sqrt(a * b); }          // contact.h:55-57
inline float combine_restitution(float a, float b) { return std::max(a, b); }        // contact.h:59-61

// Axis-aligned box (aabb.h:16-140).
struct AABB {
  Vec3f min = Vec3f::Constant(3.40282347e38f);
  Vec3f max = Vec3f::Constant(-3.40282347e38f);
  AABB() = default;
  AABB(const Vec3f& lo, const Vec3f& hi) : min(lo), max(hi) {}
  bool overlaps(const AABB& o) const {
    return (min.x() <= o.max.x() && max.x() >= o.min.x()) && (min.y() <= o.max.y() && max.y() >= o.min.y()) &&
           (min.z() <= o.max.z() && max.z() >= o.min.z());
  }
  void expand(const Vec3f& p) {
    min = min.cwiseMin(p);
    max = max.cwiseMax(p);
  }
  static AABB from_sphere(const Vec3f& c, float r) {
    const Vec3f e(r, r, r);
    return AABB(c - e, c + e);
  }
  static AABB from_oriented_box(const Vec3f& he, const Transform& t) {  // the 8 transformed corners
    AABB a;
    for (int i = 0; i < 8; ++i)
      a.expand(t.transform_point(Vec3f((i & 1) ? he[0] : -he[0], (i & 2) ? he[1] : -he[1], (i & 4) ? he[2] : -he[2])));
    return a;
  }
};

// Broadphase pair in canonical order (broadphase.h:21-30).
struct BroadPhasePair {
  int body_a = 0, body_b = 0;
  bool operator==(const BroadPhasePair& o) const { return body_a == o.body_a && body_b == o.body_b; }
};

// Sweep and prune over shape AABBs (broadphase.h:37-96, broadphase.cpp:16-118):
// endpoints sorted along x (std::sort on the first update, insertion sort
// after, keeping the previous order), a sweep from each min endpoint to the
// AABB's max, y / z overlap tests, static-static pairs skipped, then the pairs
// sorted and deduplicated.  The result is every AABB-overlapping non-static
// pair in ascending (a, b) order.
class SweepAndPrune {
 public:
  void update(const std::vector<AABB>& aabbs, const std::vector<bool>& static_mask);
  const std::vector<BroadPhasePair>& get_pairs() const { return pairs_; }

 private:
  struct Endpoint {
    float value;
    int body_index;
    bool is_min;
  };
  std::vector<Endpoint> ex_, ey_, ez_;
  std::vector<BroadPhasePair> pairs_;
  bool initialized_ = false;
};

// Narrowphase dispatcher (narrowphase.cpp:384-432): appends the contacts of
// one shape pair; true when they touch.
bool collide_shapes(const CollisionShape& a, const Transform& ta, const CollisionShape& b, const Transform& tb,
                    std::vector<ContactPoint>& contacts);
AABB shape_aabb(const CollisionShape& s, const Transform& body_transform);  // CollisionShape::compute_aabb

struct SimState {  // state.h
  std::vector<Transform> transforms;
  std::vector<Vec3f> linear_velocities, angular_velocities, forces, torques;
  void init(std::span<const Transform> t);
  void clear_forces();
  void set_linear_velocity(int i, const Vec3f& v);
  void set_angular_velocity(int i, const Vec3f& v);
  void apply_force(int i, const Vec3f& f);
  void apply_torque(int i, const Vec3f& t);
};

struct PerformancePhaseStat {
  std::string name;
  double last_ms = 0, avg_ms = 0, max_ms = 0;
  int samples = 0;
};
struct PerformanceMetric {
  std::string name;
  double value = 0;
};

// Opt-in phase timer with the reference's names and Chrome-trace export
// (performance_monitor.h:48-158).  Phases of the GPU step are filled from
// HIP-event kernel timings, mapped to the reference's phase names.
class PerformanceMonitor {
 public:
  using Clock = std::chrono::steady_clock;
  bool enabled = false;
  bool trace_enabled = false;
  int trace_frame_capacity = 120;
  PerformanceMonitor() { reset(); }
  void reset();
  void begin_frame();
  void end_frame();
  bool active() const { return frame_active_; }
  void record_phase(const std::string& name, Clock::time_point a, Clock::time_point b);
  void record_phase_ms(const std::string& name, double ms);
  void record_metric(std::string_view name, double v);
  // One GPU kernel of the current frame for the trace's GPU track: start
  // offset from the frame's first kernel and duration (ms), and its
  // algorithmic bandwidth (GB/s; 0 when not stated).
  void record_gpu_kernel(const std::string& name, double start_ms, double dur_ms, double gbps);
  std::vector<PerformancePhaseStat> phase_stats() const;
  std::vector<PerformanceMetric> last_frame_metrics() const { return last_metrics_; }
  int frame_count() const { return frame_count_; }
  double last_frame_total_ms() const { return last_total_ms_; }
  void write_trace_json(const std::string& path) const;

 private:
  struct Agg {
    double total = 0, last = 0, max = 0;
    int samples = 0;
  };
  struct Dur {
    std::string name;
    long long ts, dur;
  };
  struct GpuK {
    std::string name;
    double start_ms, dur_ms, gbps;
  };
  struct Frame {
    std::vector<Dur> dur;
    std::vector<GpuK> gpu;
    long long ts = 0, gpu_ts = 0;
    std::vector<PerformanceMetric> metrics;
  };
  long long us(Clock::time_point t) const {
    return std::chrono::duration_cast<std::chrono::microseconds>(t - epoch_).count();
  }
  Clock::time_point epoch_, frame_start_;
  bool frame_active_ = false;
  int frame_count_ = 0;
  double last_total_ms_ = 0;
  std::map<std::string, double> cur_phase_;
  std::vector<PerformanceMetric> cur_metrics_, last_metrics_;
  std::vector<Dur> cur_dur_;
  std::vector<GpuK> cur_gpu_;
  long long cur_gpu_ts_ = 0;
  std::map<std::string, Agg> agg_;
  std::deque<Frame> frames_;
};

class PhaseScope {  // PerformancePhaseScope
 public:
  PhaseScope(PerformanceMonitor* m, const char* name) : m_(m), name_(name) {
    if (m_ && m_->active()) start_ = PerformanceMonitor::Clock::now(), on_ = true;
  }
  ~PhaseScope() {
    if (on_) m_->record_phase(name_, start_, PerformanceMonitor::Clock::now());
  }

 private:
  PerformanceMonitor* m_;
  const char* name_;
  PerformanceMonitor::Clock::time_point start_;
  bool on_ = false;
};

// World: rigid-body pipeline on the host (world.cpp:30-182): integrate
// velocity, broadphase, narrowphase (sphere / box / plane), sequential-impulse
// contact solve, integrate position, clear forces.
class World {
 public:
  explicit World(const Model& model, SolverSettings settings = {});
  virtual ~World() = default;
  void step(float dt);
  void set_gravity(const Vec3f& g) { gravity_ = g; }
  const Vec3f& gravity() const { return gravity_; }
  SimState& state() { return state_; }
  const SimState& state() const { return state_; }
  const Model& model() const { return model_; }
  SolverSettings& solver_settings() { return settings_; }
  const std::vector<ContactPoint>& contacts() const { return contacts_; }
  PerformanceMonitor& performance_monitor() { return monitor_; }
  void apply_force(int body, const Vec3f& f);
  void apply_torque(int body, const Vec3f& t);

 protected:
  void step_rigid_pipeline(float dt);
  PerformanceMonitor monitor_;

 private:
  void solve_contacts(float dt);
  Model model_;
  SimState state_;
  SolverSettings settings_;
  Vec3f gravity_{0.0f, -9.81f, 0.0f};
  std::vector<ContactPoint> contacts_;
  SweepAndPrune broadphase_;
  struct ConstraintData {  // free_body_solver.h:110-120
    Vec3f r_a, r_b, tangent1, tangent2;
    float effective_mass_n = 0.0f, effective_mass_t1 = 0.0f, effective_mass_t2 = 0.0f, bias = 0.0f;
    Mat3f inv_I_a_world, inv_I_b_world;
  };
  std::vector<ConstraintData> constraint_data_;
};

// ------------------------------------------------------------ fluid
struct SPHKernels {  // sph_kernel.h:16-81 (host copies for the API)
  static float poly6(float r_sq, float h) {
    if (r_sq >= h * h) return 0.0f;
    float h2 = h * h, diff = h2 - r_sq, h9 = h2 * h2 * h2 * h2 * h;
    return (315.0f / (64.0f * PI * h9)) * diff * diff * diff;
  }
  static Vec3f spiky_grad(const Vec3f& r, float h) {
    float l = r.norm();
    if (l <= 1e-7f || l >= h) return Vec3f::Zero();
    float diff = h - l, h6 = h * h * h * h * h * h;
    return r * (-(45.0f / (PI * h6)) * diff * diff / l);
  }
  static float spiky(float l, float h) {
    if (l >= h || l < 0.0f) return 0.0f;
    float diff = h - l, h6 = h * h * h * h * h * h;
    return (15.0f / (PI * h6)) * diff * diff * diff;
  }
  static float cubic_spline(float l, float h) {
    float q = l / h;
    if (q >= 1.0f) return 0.0f;
    float sigma = 8.0f / (PI * h * h * h);
    if (q <= 0.5f) return sigma * (6.0f * (q * q * q - q * q) + 1.0f);
    float t = 1.0f - q;
    return sigma * 2.0f * t * t * t;
  }
};

struct FluidBlockDef {  // particle_state.h:52-58
  Vec3f lower, upper;
  float particle_spacing = 0.02f;
  float rest_density = 1000.0f;
  Vec3f initial_velocity;
};
std::vector<Vec3f> generate_fluid_block(const FluidBlockDef& def);

struct PBFSettings {  // pbf_solver.h:13-38 (all fields bound, C++-only ones included)
  float rest_density = 1000.0f;
  float kernel_radius = 0.1f;
  float particle_radius = 0.025f;
  int solver_iterations = 4;
  float epsilon = 100.0f;
  float xsph_viscosity = 0.01f;
  float vorticity_epsilon = 0.0f;
  float corr_delta_q = 0.3f;
  float corr_k = 0.001f;
  int corr_n = 4;
  bool use_domain_bounds = false;
  Vec3f domain_lower{-10.0f, 0.0f, -10.0f};
  Vec3f domain_upper{10.0f, 10.0f, 10.0f};
  float particle_mass(float s) const { return rest_density * s * s * s; }
  npx_pbf_params to_c() const;
};

// Device context handle (RAII over npx_ctx).
class DeviceContext {
 public:
  DeviceContext(const PBFSettings& s, int capacity, uint32_t flags);
  ~DeviceContext();
  DeviceContext(const DeviceContext&) = delete;
  DeviceContext& operator=(const DeviceContext&) = delete;
  npx_ctx* get() const { return ctx_; }
  void check(int rc) const;

 private:
  npx_ctx* ctx_ = nullptr;
};

// ParticleState (particle_state.h:16-44).  The device holds the live state
// after a step; the host vectors are a mirror refreshed lazily on access.
class ParticleState {
 public:
  void init(std::span<const Vec3f> positions, const Vec3f& v0 = Vec3f::Zero());
  void init_arrays(const float* pos3, const float* vel3, int n);
  void clear();
  int num_particles() const { return n_; }
  const std::vector<Vec3f>& positions() const;
  const std::vector<Vec3f>& velocities() const;
  const std::vector<float>& densities() const;
  const std::vector<float>& lambdas() const;
  uint64_t version() const { return version_; }
  // device binding
  void bind_device(std::shared_ptr<DeviceContext> d) {
    dev_ = std::move(d);
    host_valid_ = false;
  }
  const std::shared_ptr<DeviceContext>& device() const { return dev_; }
  bool host_valid() const { return host_valid_; }
  void sync() const;

 private:
  mutable std::vector<Vec3f> pos_, vel_;
  mutable std::vector<float> rho_, lam_;
  int n_ = 0;
  uint64_t version_ = 0;
  mutable bool host_valid_ = true;
  std::shared_ptr<DeviceContext> dev_;
};

// SpatialHashGrid (neighbor_search.h:19-99): build + query through the
// GPU grid (npx_grid_query), with the reference's bucket semantics.
class SpatialHashGrid {
 public:
  explicit SpatialHashGrid(float cell_size = 0.1f) : cell_(cell_size) {}
  void build(std::span<const Vec3f> positions) { pts_.assign(positions.begin(), positions.end()); }
  void clear() { pts_.clear(); }
  std::vector<int> query_neighbors(const Vec3f& p, float radius) const;
  std::vector<std::vector<int>> query_neighbors_batch(std::span<const Vec3f> q) const;
  std::vector<std::pair<int, int>> query_all_pairs(std::span<const Vec3f> positions, float radius) const;
  float cell_size() const { return cell_; }
  void set_cell_size(float c) { cell_ = c; }

 private:
  float cell_;
  std::vector<Vec3f> pts_;
};

class PBFSolver {  // pbf_solver.h:46-114
 public:
  explicit PBFSolver(const PBFSettings& s = {}, uint32_t flags = 0) : settings_(s), flags_(flags) {}
  void step(ParticleState& state, float dt, const Vec3f& gravity, float particle_mass);
  void step_n(ParticleState& state, float dt, const Vec3f& gravity, float particle_mass, int n);
  PBFSettings& settings() { return settings_; }
  const PBFSettings& settings() const { return settings_; }
  // device state after the last step (for coupling)
  const std::shared_ptr<DeviceContext>& context() const { return ctx_; }
  void ensure_bound(ParticleState& state);
  std::map<std::string, std::pair<double, int>> take_kernel_times(bool keep_on);
  struct KernelLaunch {
    std::string name;
    float start_ms, dur_ms;
  };
  // Timed launches in order (kernel timing must be on), before take_kernel_times.
  std::vector<KernelLaunch> kernel_trace();

 private:
  PBFSettings settings_;
  uint32_t flags_;
  std::shared_ptr<DeviceContext> ctx_;
  const ParticleState* bound_state_ = nullptr;
  uint64_t bound_version_ = ~0ull;
  npx_pbf_params last_params_{};
  bool timing_ = false;
};

struct BoundaryParticle {  // boundary.h:21-25
  Vec3f local_position;
  int body_index = -1;
  float volume = 0.0f;
};
std::vector<BoundaryParticle> sample_box_boundary(const CollisionShape& s, float spacing);
std::vector<BoundaryParticle> sample_sphere_boundary(const CollisionShape& s, float spacing);
std::vector<BoundaryParticle> sample_plane_boundary(const CollisionShape& s, float spacing, float extent = 1.0f);
std::vector<BoundaryParticle> sample_model_boundaries(const Model& model, float spacing, float plane_extent = 1.0f);
std::vector<Vec3f> boundary_world_positions(std::span<const BoundaryParticle> bp, std::span<const Transform> tf);

// FluidWorld (fluid_world.h:17-107): PBF on the GPU, Akinci boundary density
// and coupling forces on the GPU, rigid pipeline on the host.
class FluidWorld : public World {
 public:
  FluidWorld(const Model& model, const std::vector<FluidBlockDef>& blocks = {}, SolverSettings ss = {},
             PBFSettings pbf = {}, float boundary_extent = 1.0f, uint32_t flags = 0);
  void step(float dt);
  ParticleState& fluid_state() { return state_; }
  PBFSettings& pbf_settings() { return solver_.settings(); }
  int num_particles() const { return state_.num_particles(); }
  int num_boundary_particles() const { return (int)boundary_.size(); }
  // psi volumes are computed on the device the first time they are needed
  const std::vector<BoundaryParticle>& boundary_particles() {
    if (!boundary_.empty() && !boundary_uploaded_) upload_boundary();
    return boundary_;
  }
  float particle_mass() const { return particle_mass_; }

 private:
  void upload_boundary();
  std::vector<float> poses7() const;
  ParticleState state_;
  PBFSolver solver_;
  std::vector<BoundaryParticle> boundary_;
  float particle_mass_ = 0.0f;
  float particle_spacing_ = 0.02f;
  bool boundary_uploaded_ = false;
  const ParticleState* boundary_state_ = nullptr;
};

}  // namespace novaphy
