// fluid.cpp -- host side of the fluid path: ParticleState mirror, PBFSolver,
// SpatialHashGrid, boundary sampling and FluidWorld, all computing through
// libnpx_hip's C ABI.  Setup-time geometry (lattice blocks, boundary
// sampling) stays on the host like in the reference.
#include <algorithm>
#include <cstring>

#include "npx_host.hpp"

namespace novaphy {

static void throw_npx(int rc, const char* msg) {
  std::string m = std::string("libnpx_hip: ") + (msg ? msg : "error");
  if (rc == NPX_EINVAL) throw std::invalid_argument(m);
  throw std::runtime_error(m);
}

npx_pbf_params PBFSettings::to_c() const {
  npx_pbf_params p{};
  p.rest_density = rest_density;
  p.kernel_radius = kernel_radius;
  p.particle_radius = particle_radius;
  p.solver_iterations = solver_iterations;
  p.epsilon = epsilon;
  p.xsph_viscosity = xsph_viscosity;
  p.vorticity_epsilon = vorticity_epsilon;
  p.corr_delta_q = corr_delta_q;
  p.corr_k = corr_k;
  p.corr_n = corr_n;
  p.use_domain_bounds = use_domain_bounds ? 1 : 0;
  for (int d = 0; d < 3; ++d) {
    p.domain_lower[d] = domain_lower[d];
    p.domain_upper[d] = domain_upper[d];
  }
  return p;
}

DeviceContext::DeviceContext(const PBFSettings& s, int capacity, uint32_t flags) {
  npx_pbf_params p = s.to_c();
  int rc = npx_create(&p, capacity, flags, &ctx_);
  if (rc != NPX_OK) throw_npx(rc, npx_last_error(nullptr));
}
DeviceContext::~DeviceContext() {
  if (ctx_) npx_destroy(ctx_);
}
void DeviceContext::check(int rc) const {
  if (rc != NPX_OK) throw_npx(rc, npx_last_error(ctx_));
}

// particle_state.cpp:31-58
std::vector<Vec3f> generate_fluid_block(const FluidBlockDef& def) {
  std::vector<Vec3f> out;
  const float s = def.particle_spacing;
  if (s <= 0.0f) return out;
  int n[3];
  for (int d = 0; d < 3; ++d)
    n[d] = std::max(1, static_cast<int>(std::floor((def.upper[d] - def.lower[d]) / s)) + 1);
  out.reserve((size_t)n[0] * n[1] * n[2]);
  for (int iz = 0; iz < n[2]; ++iz)
    for (int iy = 0; iy < n[1]; ++iy)
      for (int ix = 0; ix < n[0]; ++ix)
        out.emplace_back(def.lower[0] + ix * s, def.lower[1] + iy * s, def.lower[2] + iz * s);
  return out;
}

// ------------------------------------------------------------ ParticleState
void ParticleState::init(std::span<const Vec3f> positions, const Vec3f& v0) {
  n_ = (int)positions.size();
  pos_.assign(positions.begin(), positions.end());
  vel_.assign(n_, v0);
  rho_.assign(n_, 0.0f);
  lam_.assign(n_, 0.0f);
  ++version_;
  host_valid_ = true;
}

void ParticleState::init_arrays(const float* pos3, const float* vel3, int n) {
  n_ = n;
  pos_.resize(n);
  vel_.resize(n);
  for (int i = 0; i < n; ++i) {
    pos_[i] = Vec3f(pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2]);
    vel_[i] = vel3 ? Vec3f(vel3[3 * i], vel3[3 * i + 1], vel3[3 * i + 2]) : Vec3f::Zero();
  }
  rho_.assign(n, 0.0f);
  lam_.assign(n, 0.0f);
  ++version_;
  host_valid_ = true;
}

void ParticleState::clear() {
  n_ = 0;
  pos_.clear();
  vel_.clear();
  rho_.clear();
  lam_.clear();
  ++version_;
  host_valid_ = true;
}

void ParticleState::sync() const {
  if (host_valid_ || !dev_ || n_ == 0) {
    host_valid_ = true;
    return;
  }
  std::vector<float> p(3 * (size_t)n_), v(3 * (size_t)n_);
  rho_.resize(n_);
  lam_.resize(n_);
  dev_->check(npx_get_particles(dev_->get(), p.data(), v.data(), rho_.data(), lam_.data()));
  pos_.resize(n_);
  vel_.resize(n_);
  for (int i = 0; i < n_; ++i) {
    pos_[i] = Vec3f(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
    vel_[i] = Vec3f(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
  }
  host_valid_ = true;
}

const std::vector<Vec3f>& ParticleState::positions() const {
  sync();
  return pos_;
}
const std::vector<Vec3f>& ParticleState::velocities() const {
  sync();
  return vel_;
}
const std::vector<float>& ParticleState::densities() const {
  sync();
  return rho_;
}
const std::vector<float>& ParticleState::lambdas() const {
  sync();
  return lam_;
}

// ------------------------------------------------------------ PBFSolver
void PBFSolver::ensure_bound(ParticleState& state) {
  const int n = state.num_particles();
  const npx_pbf_params p = settings_.to_c();
  const bool fresh = !ctx_ || bound_state_ != &state || bound_version_ != state.version() ||
                     state.device() != ctx_;
  if (!ctx_) {
    ctx_ = std::make_shared<DeviceContext>(settings_, std::max(n, 1), flags_);
    last_params_ = p;
  }
  if (std::memcmp(&p, &last_params_, sizeof(p)) != 0) {
    ctx_->check(npx_set_params(ctx_->get(), &p));
    last_params_ = p;
  }
  if (fresh) {
    state.sync();  // host vectors must hold the newest data before uploading
    const auto& pos = state.positions();
    const auto& vel = state.velocities();
    std::vector<float> pf(3 * (size_t)n), vf(3 * (size_t)n);
    for (int i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d) {
        pf[3 * i + d] = pos[i][d];
        vf[3 * i + d] = vel[i][d];
      }
    ctx_->check(npx_set_particles(ctx_->get(), pf.data(), vf.data(), n));
    state.bind_device(ctx_);
    bound_state_ = &state;
    bound_version_ = state.version();
  }
}

void PBFSolver::step_n(ParticleState& state, float dt, const Vec3f& gravity, float particle_mass, int nsteps) {
  if (state.num_particles() == 0) return;  // pbf_solver.cpp:21
  ensure_bound(state);
  const float g[3] = {gravity[0], gravity[1], gravity[2]};
  ctx_->check(npx_step_n(ctx_->get(), dt, g, particle_mass, nsteps));
  state.bind_device(ctx_);  // device now holds the newest state
}

void PBFSolver::step(ParticleState& state, float dt, const Vec3f& gravity, float particle_mass) {
  step_n(state, dt, gravity, particle_mass, 1);
}

std::vector<PBFSolver::KernelLaunch> PBFSolver::kernel_trace() {
  std::vector<KernelLaunch> out;
  if (!ctx_ || !timing_) return out;
  int cnt = 0;
  ctx_->check(npx_kernel_trace(ctx_->get(), nullptr, 0, nullptr, nullptr, 0, &cnt));
  if (cnt <= 0) return out;
  std::vector<char> names((size_t)cnt * 40 + 64);
  std::vector<float> st(cnt), du(cnt);
  ctx_->check(npx_kernel_trace(ctx_->get(), names.data(), (int32_t)names.size(), st.data(), du.data(), cnt, &cnt));
  const char* p = names.data();
  for (int k = 0; k < cnt && *p; ++k) {
    const char* e = std::strchr(p, '\n');
    std::string nm = e ? std::string(p, e - p) : std::string(p);
    out.push_back({nm, st[k], du[k]});
    p = e ? e + 1 : p + nm.size();
  }
  return out;
}

std::map<std::string, std::pair<double, int>> PBFSolver::take_kernel_times(bool keep_on) {
  std::map<std::string, std::pair<double, int>> out;
  if (!ctx_) return out;
  if (timing_) {
    int cnt = 0;
    char names[8192];
    float ms[128];
    int32_t nl[128];
    ctx_->check(npx_kernel_times(ctx_->get(), names, sizeof(names), ms, nl, 128, &cnt));
    const char* p = names;
    for (int k = 0; k < cnt && k < 128 && *p; ++k) {
      const char* e = std::strchr(p, '\n');
      std::string nm = e ? std::string(p, e - p) : std::string(p);
      out[nm] = {ms[k], nl[k]};
      p = e ? e + 1 : p + nm.size();
    }
  }
  ctx_->check(npx_set_kernel_timing(ctx_->get(), keep_on ? 1 : 0));
  timing_ = keep_on;
  return out;
}

// ------------------------------------------------------------ SpatialHashGrid
std::vector<std::vector<int>> SpatialHashGrid::query_neighbors_batch(std::span<const Vec3f> q) const {
  std::vector<float> pf(3 * pts_.size()), qf(3 * q.size());
  for (size_t i = 0; i < pts_.size(); ++i)
    for (int d = 0; d < 3; ++d) pf[3 * i + d] = pts_[i][d];
  for (size_t i = 0; i < q.size(); ++i)
    for (int d = 0; d < 3; ++d) qf[3 * i + d] = q[i][d];
  std::vector<int32_t> counts(q.size());
  int64_t total = 0;
  int rc = npx_grid_query(pf.data(), (int)pts_.size(), cell_, qf.data(), (int)q.size(), counts.data(), nullptr, 0,
                          &total);
  if (rc != NPX_OK) throw_npx(rc, npx_last_error(nullptr));
  std::vector<int32_t> lists(std::max<int64_t>(total, 1));
  rc = npx_grid_query(pf.data(), (int)pts_.size(), cell_, qf.data(), (int)q.size(), counts.data(), lists.data(),
                      total, &total);
  if (rc != NPX_OK) throw_npx(rc, npx_last_error(nullptr));
  std::vector<std::vector<int>> out(q.size());
  int64_t o = 0;
  for (size_t i = 0; i < q.size(); ++i) {
    out[i].assign(lists.begin() + o, lists.begin() + o + counts[i]);
    o += counts[i];
  }
  return out;
}

std::vector<int> SpatialHashGrid::query_neighbors(const Vec3f& p, float radius) const {
  (void)radius;  // the reference ignores it too (neighbor_search.cpp:30)
  if (pts_.empty()) return {};
  Vec3f q[1] = {p};
  return query_neighbors_batch(std::span<const Vec3f>(q, 1))[0];
}

std::vector<std::pair<int, int>> SpatialHashGrid::query_all_pairs(std::span<const Vec3f> positions,
                                                                  float radius) const {
  // neighbor_search.cpp:57-87: the built grid's candidates of each given
  // point, j > i, with |positions[i] - positions[j]|^2 < r^2 (both from the
  // given positions, as the reference; a candidate index past them, which the
  // reference would read out of bounds, is skipped)
  std::vector<std::pair<int, int>> pairs;
  if (positions.empty()) return pairs;
  const float r2 = radius * radius;
  auto lists = query_neighbors_batch(positions);
  for (size_t i = 0; i < positions.size(); ++i)
    for (int j : lists[i]) {
      if (j <= (int)i || j >= (int)positions.size()) continue;
      if ((positions[i] - positions[(size_t)j]).squaredNorm() < r2) pairs.emplace_back((int)i, j);
    }
  return pairs;
}

// ------------------------------------------------------------ boundary sampling
// boundary.cpp:15-150 (setup geometry; host side like the reference)
std::vector<BoundaryParticle> sample_box_boundary(const CollisionShape& s, float spacing) {
  std::vector<BoundaryParticle> out;
  if (s.type != ShapeType::Box) return out;
  const Vec3f& he = s.half_extents;
  for (int axis = 0; axis < 3; ++axis) {
    const int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
    const int n1 = std::max(1, static_cast<int>(std::floor(2.0f * he[a1] / spacing)) + 1);
    const int n2 = std::max(1, static_cast<int>(std::floor(2.0f * he[a2] / spacing)) + 1);
    for (int sign = -1; sign <= 1; sign += 2)
      for (int i1 = 0; i1 < n1; ++i1)
        for (int i2 = 0; i2 < n2; ++i2) {
          Vec3f lp;
          lp[axis] = sign * he[axis];
          lp[a1] = -he[a1] + i1 * spacing;
          lp[a2] = -he[a2] + i2 * spacing;
          BoundaryParticle b;
          b.local_position = s.local_transform.transform_point(lp);
          b.body_index = s.body_index;
          out.push_back(b);
        }
  }
  return out;
}

std::vector<BoundaryParticle> sample_sphere_boundary(const CollisionShape& s, float spacing) {
  std::vector<BoundaryParticle> out;
  if (s.type != ShapeType::Sphere) return out;
  const float r = s.radius;
  const float pi = 3.14159265358979323846f;
  const float area = 4.0f * pi * r * r;
  const int n = std::max(12, static_cast<int>(area / (spacing * spacing)));
  const float golden_ratio = (1.0f + std::sqrt(5.0f)) / 2.0f;
  const float golden_angle = 2.0f * pi / (golden_ratio * golden_ratio);
  for (int i = 0; i < n; ++i) {  // Fibonacci sphere
    const float theta = std::acos(1.0f - 2.0f * (i + 0.5f) / n);
    const float phi = golden_angle * i;
    const Vec3f lp(r * std::sin(theta) * std::cos(phi), r * std::sin(theta) * std::sin(phi), r * std::cos(theta));
    BoundaryParticle b;
    b.local_position = s.local_transform.transform_point(lp);
    b.body_index = s.body_index;
    out.push_back(b);
  }
  return out;
}

std::vector<BoundaryParticle> sample_plane_boundary(const CollisionShape& s, float spacing, float extent) {
  std::vector<BoundaryParticle> out;
  if (s.type != ShapeType::Plane) return out;
  const Vec3f& nrm = s.normal;
  const Vec3f t1 = (std::abs(nrm[1]) < 0.9f ? Vec3f(0, 1, 0).cross(nrm) : Vec3f(1, 0, 0).cross(nrm)).normalized();
  const Vec3f t2 = nrm.cross(t1).normalized();
  const int n = static_cast<int>(std::floor(2.0f * extent / spacing)) + 1;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const float a = -extent + i * spacing, c = -extent + j * spacing;
      BoundaryParticle b;
      b.local_position = (nrm * s.offset + t1 * a) + t2 * c;
      b.body_index = s.body_index;
      out.push_back(b);
    }
  return out;
}

std::vector<BoundaryParticle> sample_model_boundaries(const Model& model, float spacing, float plane_extent) {
  std::vector<BoundaryParticle> all;
  for (const auto& s : model.shapes) {
    std::vector<BoundaryParticle> part;
    if (s.type == ShapeType::Box)
      part = sample_box_boundary(s, spacing);
    else if (s.type == ShapeType::Sphere)
      part = sample_sphere_boundary(s, spacing);
    else
      part = sample_plane_boundary(s, spacing, plane_extent);
    all.insert(all.end(), part.begin(), part.end());
  }
  return all;
}

std::vector<Vec3f> boundary_world_positions(std::span<const BoundaryParticle> bp, std::span<const Transform> tf) {
  std::vector<Vec3f> out(bp.size());
  for (size_t i = 0; i < bp.size(); ++i) {
    const int b = bp[i].body_index;
    out[i] = (b >= 0 && b < (int)tf.size()) ? tf[b].transform_point(bp[i].local_position) : bp[i].local_position;
  }
  return out;
}

// ------------------------------------------------------------ FluidWorld
FluidWorld::FluidWorld(const Model& model, const std::vector<FluidBlockDef>& blocks, SolverSettings ss,
                       PBFSettings pbf, float boundary_extent, uint32_t flags)
    : World(model, ss), solver_(pbf, flags) {  // fluid_world.cpp:11-44
  std::vector<Vec3f> all;
  Vec3f v0;
  for (const auto& b : blocks) {
    auto p = generate_fluid_block(b);
    all.insert(all.end(), p.begin(), p.end());
    v0 = b.initial_velocity;  // initial velocity and spacing come from the last block
    particle_spacing_ = b.particle_spacing;
  }
  if (!all.empty()) {
    state_.init(all, v0);
    particle_mass_ = pbf.particle_mass(particle_spacing_);
  }
  boundary_ = sample_model_boundaries(model, particle_spacing_, boundary_extent);
}

std::vector<float> FluidWorld::poses7() const {
  const auto& tf = state().transforms;
  std::vector<float> p(7 * std::max<size_t>(tf.size(), 1));
  for (size_t k = 0; k < tf.size(); ++k) {
    p[7 * k] = tf[k].position[0];
    p[7 * k + 1] = tf[k].position[1];
    p[7 * k + 2] = tf[k].position[2];
    p[7 * k + 3] = tf[k].rotation.w();
    p[7 * k + 4] = tf[k].rotation.x();
    p[7 * k + 5] = tf[k].rotation.y();
    p[7 * k + 6] = tf[k].rotation.z();
  }
  return p;
}

// Boundary particles and their psi volumes live on the solver's context;
// psi is computed there at the initial poses (compute_boundary_volumes at
// construction, fluid_world.cpp:39-43).
void FluidWorld::upload_boundary() {
  solver_.ensure_bound(state_);
  if (boundary_uploaded_ && boundary_state_ == &state_) return;
  const int nb = (int)boundary_.size();
  std::vector<float> local(3 * (size_t)nb);
  std::vector<int32_t> body(nb);
  for (int b = 0; b < nb; ++b) {
    for (int d = 0; d < 3; ++d) local[3 * b + d] = boundary_[b].local_position[d];
    body[b] = boundary_[b].body_index;
  }
  npx_ctx* c = solver_.context()->get();
  const int nbody = model().num_bodies();
  solver_.context()->check(npx_set_boundary(c, local.data(), body.data(), nullptr, nb, nbody));
  std::vector<float> psi(nb);
  // volumes at the model's initial transforms
  std::vector<float> p0(7 * std::max(nbody, 1));
  for (int k = 0; k < nbody; ++k) {
    const Transform& t = model().initial_transforms[k];
    const float v[7] = {t.position[0], t.position[1], t.position[2], t.rotation.w(), t.rotation.x(),
                        t.rotation.y(), t.rotation.z()};
    std::copy(v, v + 7, p0.begin() + 7 * k);
  }
  solver_.context()->check(npx_compute_boundary_volumes(c, p0.data(), nbody, psi.data()));
  for (int b = 0; b < nb; ++b) boundary_[b].volume = psi[b];
  boundary_uploaded_ = true;
  boundary_state_ = &state_;
}

void FluidWorld::step(float dt) {  // fluid_world.cpp:46-99
  monitor_.begin_frame();
  {
    PhaseScope total(&monitor_, "fluid.total");
    const int n_fluid = state_.num_particles();
    const int nb = (int)boundary_.size();
    const bool profile = monitor_.active();
    if (n_fluid > 0) {
      if (profile) solver_.ensure_bound(state_), solver_.take_kernel_times(true);
      std::vector<float> poses;
      if (nb > 0) {
        PhaseScope p(&monitor_, "fluid.boundary.positions");
        upload_boundary();
        poses = poses7();
        solver_.context()->check(
            npx_boundary_world_positions(solver_.context()->get(), poses.data(), model().num_bodies(), nullptr));
      }
      {
        PhaseScope p(&monitor_, "fluid.pbf.total");
        solver_.step(state_, dt, gravity(), particle_mass_);
        if (profile) solver_.context()->check(npx_synchronize(solver_.context()->get()));
      }
      if (nb > 0) {
        PhaseScope p(&monitor_, "fluid.boundary_density");
        solver_.context()->check(npx_apply_boundary_density(solver_.context()->get()));
      }
      if (nb > 0) {
        PhaseScope p(&monitor_, "fluid.coupling.total");
        const int nbody = model().num_bodies();
        std::vector<uint8_t> st(std::max(nbody, 1));
        for (int k = 0; k < nbody; ++k) st[k] = model().bodies[k].is_static() ? 1 : 0;
        std::vector<float> f(3 * std::max(nbody, 1)), t(3 * std::max(nbody, 1));
        solver_.context()->check(
            npx_coupling_forces(solver_.context()->get(), poses.data(), st.data(), nbody, f.data(), t.data()));
        for (int k = 0; k < nbody; ++k) {
          if (st[k]) continue;
          const Vec3f F(f[3 * k], f[3 * k + 1], f[3 * k + 2]), T(t[3 * k], t[3 * k + 1], t[3 * k + 2]);
          if (F.squaredNorm() == 0.0f && T.squaredNorm() == 0.0f) continue;
          World::apply_force(k, F);
          World::apply_torque(k, T);
        }
      }
      state_.bind_device(solver_.context());
      if (profile) {
        // GPU kernel time under the reference's phase names.  Fused passes
        // are reported under the first phase they cover: density + lambda
        // under compute_density, delta-p + apply (+ the velocity update on
        // the last iteration) under compute_delta_position, XSPH + vorticity
        // curl under xsph_viscosity, vorticity force + commit under
        // commit_positions.
        static const std::pair<const char*, const char*> map[] = {
            {"predict_bin", "fluid.pbf.predict"},
            {"fill_counts", "fluid.pbf.grid_build"},
            {"scan_reduce", "fluid.pbf.grid_build"},
            {"scan_bsum", "fluid.pbf.grid_build"},
            {"scan_down", "fluid.pbf.grid_build"},
            {"scatter", "fluid.pbf.grid_build"},
            {"rank_sort", "fluid.pbf.neighbor_query"},
            {"rank", "fluid.pbf.neighbor_query"},
            {"rank_merge", "fluid.pbf.neighbor_query"},
            {"huge_hist", "fluid.pbf.neighbor_query"},
            {"huge_scan", "fluid.pbf.neighbor_query"},
            {"huge_place", "fluid.pbf.neighbor_query"},
            {"gather", "fluid.pbf.neighbor_query"},
            {"pieces", "fluid.pbf.neighbor_query"},
            {"clear_cmax", "fluid.pbf.neighbor_query"},
            {"dense_reset", "fluid.pbf.neighbor_query"},
            {"density_lambda_build", "fluid.pbf.iteration.compute_density"},
            {"density_lambda_build_lists", "fluid.pbf.iteration.compute_density"},
            {"density_lambda", "fluid.pbf.iteration.compute_density"},
            {"density_lambda_build_dense", "fluid.pbf.iteration.compute_density"},
            {"density_lambda_dense", "fluid.pbf.iteration.compute_density"},
            {"delta", "fluid.pbf.iteration.compute_delta_position"},
            {"delta_dense", "fluid.pbf.iteration.compute_delta_position"},
            {"omega_xsph", "fluid.pbf.xsph_viscosity"},
            {"omega_xsph_dense", "fluid.pbf.xsph_viscosity"},
            {"vort_commit", "fluid.pbf.commit_positions"},
            {"vort_commit_dense", "fluid.pbf.commit_positions"},
            {"run_fill", "fluid.pbf.commit_positions"},
        };
        // algorithmic bytes per particle of each pass (FP32/int32, 12 B per
        // 3-vector), for the GB/s counters of the trace's GPU track.  The
        // build pass is two launches of which one does the work (with or
        // without skin lists), so it has no per-launch rate.
        const bool vort = solver_.settings().vorticity_epsilon != 0.0f;
        const std::map<std::string, double> bytes = {
            {"predict_bin", 40}, {"gather", 56}, {"density_lambda", 20},
            {"delta", 28}, {"omega_xsph", vort ? 36.0 : 24.0}, {"vort_commit", vort ? 64.0 : 28.0}};
        for (const auto& k : solver_.kernel_trace()) {
          auto b = bytes.find(k.name);
          const double gbps = (b != bytes.end() && k.dur_ms > 0.0f)
                                  ? b->second * (double)n_fluid / ((double)k.dur_ms * 1e-3) / 1e9
                                  : 0.0;
          monitor_.record_gpu_kernel(k.name, k.start_ms, k.dur_ms, gbps);
        }
        auto kt = solver_.take_kernel_times(false);
        for (const auto& [k, ph] : map) {
          auto it = kt.find(k);
          if (it != kt.end()) monitor_.record_phase_ms(ph, it->second.first);
        }
        // the boundary and coupling kernels (no counterpart phase in the
        // reference) under their own names
        for (const auto& [k, v] : kt) {
          bool mapped = false;
          for (const auto& m : map) mapped = mapped || k == m.first;
          if (!mapped) monitor_.record_phase_ms("fluid.gpu." + k, v.first);
        }
      }
    }
    {
      PhaseScope p(&monitor_, "fluid.rigid_step");
      step_rigid_pipeline(dt);
    }
    monitor_.record_metric("fluid_particles", (double)state_.num_particles());
    monitor_.record_metric("boundary_particles", (double)boundary_.size());
    monitor_.record_metric("pbf_solver_iterations", (double)solver_.settings().solver_iterations);
    monitor_.record_metric("kernel_radius", (double)solver_.settings().kernel_radius);
  }
  monitor_.end_frame();
}

}  // namespace novaphy
