// rigid.cpp -- math helpers, model factories and the host rigid-body pipeline.
//
// The rigid pipeline stays on the host, restated from the reference's
// World::step_rigid_pipeline (src/sim/world.cpp:30-182): semi-implicit Euler
// with 0.999 angular damping (src/dynamics/integrator.cpp), sweep-and-prune
// over shape AABBs (src/collision/broadphase.cpp), the narrowphase of
// src/collision/narrowphase.cpp (sphere / box / plane pairs, 15-axis SAT for
// box-box), the warm-started PGS contact solver of
// src/dynamics/free_body_solver.cpp, quaternion integration and force
// clearing.  Sleeping (off by default in the reference) is not modelled.
#include <algorithm>

#include "npx_host.hpp"

namespace novaphy {

// Inverse by cofactors, in the order of Eigen's 3x3 inverse
// (compute_inverse_size3_helper): det = c00 m00 + (c10 m10 + c20 m20), each
// cofactor a 2x2 determinant, every entry cofactor * (1 / det).
Mat3f Mat3f::inverse() const {
  auto cof = [&](int i, int j) {
    const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    return m[i1][j1] * m[i2][j2] - m[i1][j2] * m[i2][j1];
  };
  const float c00 = cof(0, 0), c10 = cof(1, 0), c20 = cof(2, 0);
  const float det = c00 * m[0][0] + (c10 * m[1][0] + c20 * m[2][0]);
  const float inv = 1.0f / det;
  Mat3f r;
  r.m[0][0] = c00 * inv;
  r.m[0][1] = c10 * inv;
  r.m[0][2] = c20 * inv;
  r.m[1][0] = cof(0, 1) * inv;
  r.m[1][1] = cof(1, 1) * inv;
  r.m[1][2] = cof(2, 1) * inv;
  r.m[2][0] = cof(0, 2) * inv;
  r.m[2][1] = cof(1, 2) * inv;
  r.m[2][2] = cof(2, 2) * inv;
  return r;
}

// Rotation matrix of a unit quaternion, in the order of Eigen's
// Quaternion::toRotationMatrix (products of the doubled components).
Mat3f Quatf::matrix() const {
  Mat3f r;
  const float tx = 2.0f * x_, ty = 2.0f * y_, tz = 2.0f * z_;
  const float twx = tx * w_, twy = ty * w_, twz = tz * w_;
  const float txx = tx * x_, txy = ty * x_, txz = tz * x_;
  const float tyy = ty * y_, tyz = tz * y_, tzz = tz * z_;
  r.m[0][0] = 1.0f - (tyy + tzz);
  r.m[0][1] = txy - twz;
  r.m[0][2] = txz + twy;
  r.m[1][0] = txy + twz;
  r.m[1][1] = 1.0f - (txx + tzz);
  r.m[1][2] = tyz - twx;
  r.m[2][0] = txz - twy;
  r.m[2][1] = tyz + twx;
  r.m[2][2] = 1.0f - (txx + tyy);
  return r;
}

RigidBody RigidBody::from_box(float m, const Vec3f& he) {  // body.h:63-78
  RigidBody b;
  b.mass = m;
  const float w = 2.0f * he[0], h = 2.0f * he[1], d = 2.0f * he[2], c = m / 12.0f;
  b.inertia = Mat3f::Zero();
  b.inertia(0, 0) = c * (h * h + d * d);
  b.inertia(1, 1) = c * (w * w + d * d);
  b.inertia(2, 2) = c * (w * w + h * h);
  return b;
}

RigidBody RigidBody::from_sphere(float m, float radius) {  // body.h:80-90
  RigidBody b;
  b.mass = m;
  b.inertia = Mat3f::Identity() * ((2.0f / 5.0f) * m * radius * radius);
  return b;
}

RigidBody RigidBody::make_static() {
  RigidBody b;
  b.mass = 0.0f;
  b.inertia = Mat3f::Zero();
  return b;
}

CollisionShape CollisionShape::make_box(const Vec3f& he, int body, const Transform& local, float fr, float re) {
  CollisionShape s;
  s.type = ShapeType::Box;
  s.half_extents = he;
  s.body_index = body;
  s.local_transform = local;
  s.friction = fr;
  s.restitution = re;
  return s;
}

CollisionShape CollisionShape::make_sphere(float r, int body, const Transform& local, float fr, float re) {
  CollisionShape s;
  s.type = ShapeType::Sphere;
  s.radius = r;
  s.body_index = body;
  s.local_transform = local;
  s.friction = fr;
  s.restitution = re;
  return s;
}

CollisionShape CollisionShape::make_plane(const Vec3f& n, float offset, float fr, float re) {
  CollisionShape s;
  s.type = ShapeType::Plane;
  s.normal = n.normalized();
  s.offset = offset;
  s.body_index = -1;
  s.friction = fr;
  s.restitution = re;
  return s;
}

// ------------------------------------------------------------- SimState
void SimState::init(std::span<const Transform> t) {
  transforms.assign(t.begin(), t.end());
  const size_t n = t.size();
  linear_velocities.assign(n, Vec3f::Zero());
  angular_velocities.assign(n, Vec3f::Zero());
  forces.assign(n, Vec3f::Zero());
  torques.assign(n, Vec3f::Zero());
}
void SimState::clear_forces() {
  std::fill(forces.begin(), forces.end(), Vec3f::Zero());
  std::fill(torques.begin(), torques.end(), Vec3f::Zero());
}
static void check_body(const SimState& s, int i) {  // setters name a body that must exist
  if (i < 0 || i >= (int)s.transforms.size()) throw std::out_of_range("body index out of range");
}
void SimState::set_linear_velocity(int i, const Vec3f& v) {
  check_body(*this, i);
  linear_velocities[i] = v;
}
void SimState::set_angular_velocity(int i, const Vec3f& v) {
  check_body(*this, i);
  angular_velocities[i] = v;
}
void SimState::apply_force(int i, const Vec3f& f) {  // state.cpp:63-67: out-of-range indices are ignored
  if (i >= 0 && i < (int)forces.size()) forces[i] += f;
}
void SimState::apply_torque(int i, const Vec3f& t) {  // state.cpp:74-78
  if (i >= 0 && i < (int)torques.size()) torques[i] += t;
}

// ------------------------------------------------------------- World
World::World(const Model& model, SolverSettings settings) : model_(model), settings_(settings) {
  state_.init(model_.initial_transforms);
}

void World::step(float dt) {
  monitor_.begin_frame();
  step_rigid_pipeline(dt);
  monitor_.end_frame();
}

void World::apply_force(int body, const Vec3f& f) { state_.apply_force(body, f); }
void World::apply_torque(int body, const Vec3f& t) { state_.apply_torque(body, t); }


// ------------------------------------------------------------- broadphase
void SweepAndPrune::update(const std::vector<AABB>& aabbs, const std::vector<bool>& static_mask) {
  const int n = (int)aabbs.size();
  ex_.resize(2 * (size_t)n);
  ey_.resize(2 * (size_t)n);
  ez_.resize(2 * (size_t)n);
  for (int i = 0; i < n; ++i) {
    ex_[2 * i] = {aabbs[i].min.x(), i, true};
    ex_[2 * i + 1] = {aabbs[i].max.x(), i, false};
    ey_[2 * i] = {aabbs[i].min.y(), i, true};
    ey_[2 * i + 1] = {aabbs[i].max.y(), i, false};
    ez_[2 * i] = {aabbs[i].min.z(), i, true};
    ez_[2 * i + 1] = {aabbs[i].max.z(), i, false};
  }
  auto by_value = [](const Endpoint& a, const Endpoint& b) { return a.value < b.value; };
  auto insertion = [](std::vector<Endpoint>& e) {  // temporal coherence: nearly sorted after the first frame
    for (size_t i = 1; i < e.size(); ++i) {
      const Endpoint k = e[i];
      size_t j = i;
      while (j > 0 && e[j - 1].value > k.value) {
        e[j] = e[j - 1];
        --j;
      }
      e[j] = k;
    }
  };
  for (auto* e : {&ex_, &ey_, &ez_}) {
    if (initialized_)
      insertion(*e);
    else
      std::sort(e->begin(), e->end(), by_value);
  }
  initialized_ = true;
  pairs_.clear();
  for (size_t i = 0; i < ex_.size(); ++i) {
    if (!ex_[i].is_min) continue;
    const int a = ex_[i].body_index;
    const float max_a = aabbs[a].max.x();
    for (size_t j = i + 1; j < ex_.size(); ++j) {
      if (ex_[j].value > max_a) break;
      if (!ex_[j].is_min) continue;
      const int b = ex_[j].body_index;
      if (a == b || (static_mask[a] && static_mask[b])) continue;
      const AABB& A = aabbs[a];
      const AABB& B = aabbs[b];
      if (A.min.y() <= B.max.y() && A.max.y() >= B.min.y() && A.min.z() <= B.max.z() && A.max.z() >= B.min.z())
        pairs_.push_back({std::min(a, b), std::max(a, b)});
    }
  }
  std::sort(pairs_.begin(), pairs_.end(), [](const BroadPhasePair& p, const BroadPhasePair& q) {
    return p.body_a < q.body_a || (p.body_a == q.body_a && p.body_b < q.body_b);
  });
  pairs_.erase(std::unique(pairs_.begin(), pairs_.end()), pairs_.end());
}

AABB shape_aabb(const CollisionShape& s, const Transform& body_transform) {  // shape.h:62-75
  const Transform w = body_transform * s.local_transform;
  switch (s.type) {
    case ShapeType::Box:
      return AABB::from_oriented_box(s.half_extents, w);
    case ShapeType::Sphere:
      return AABB::from_sphere(w.position, s.radius);
    case ShapeType::Plane:
      return AABB(Vec3f::Constant(-1e6f), Vec3f::Constant(1e6f));
  }
  return AABB();
}

// ------------------------------------------------------------- narrowphase
namespace {

ContactPoint contact(const Vec3f& pos, const Vec3f& n, float pen, int ba, int bb, float fr, float re) {
  ContactPoint c;
  c.position = pos;
  c.normal = n;
  c.penetration = pen;
  c.body_a = ba;
  c.body_b = bb;
  c.friction = fr;
  c.restitution = re;
  return c;
}

// narrowphase.cpp:25-58
bool sphere_sphere(const CollisionShape& a, const Transform& ta, const CollisionShape& b, const Transform& tb,
                   std::vector<ContactPoint>& out) {
  const Vec3f ca = ta.transform_point(a.local_transform.position);
  const Vec3f cb = tb.transform_point(b.local_transform.position);
  const float ra = a.radius, rb = b.radius;
  const Vec3f diff = cb - ca;
  const float dist_sq = diff.squaredNorm();
  const float rsum = ra + rb;
  if (dist_sq > rsum * rsum) return false;
  const float dist = std::sqrt(dist_sq);
  const Vec3f n = dist < 1e-6f ? Vec3f(0, 1, 0) : diff / dist;
  const float pen = rsum - dist;
  out.push_back(contact(ca + n * (ra - pen * 0.5f), n, pen, a.body_index, b.body_index,
                        combine_friction(a.friction, b.friction), combine_restitution(a.restitution, b.restitution)));
  return true;
}

// narrowphase.cpp:69-94: the plane is body A, the normal points from the plane toward the sphere
bool sphere_plane(const CollisionShape& s, const Transform& ts, const CollisionShape& p,
                  std::vector<ContactPoint>& out) {
  const Vec3f center = ts.transform_point(s.local_transform.position);
  const float radius = s.radius;
  const float distance = p.normal.dot(center) - p.offset;
  if (distance > radius) return false;
  out.push_back(contact(center - p.normal * distance, p.normal, radius - distance, p.body_index, s.body_index,
                        combine_friction(s.friction, p.friction), combine_restitution(s.restitution, p.restitution)));
  return true;
}

// narrowphase.cpp:107-165: closest point of the box to the sphere centre in the
// box frame; a centre inside the box takes the nearest face's normal
bool box_sphere(const CollisionShape& bs, const Transform& tb, const CollisionShape& ss, const Transform& ts,
                std::vector<ContactPoint>& out) {
  const Transform box_world = tb * bs.local_transform;
  const Transform box_inv = box_world.inverse();
  const Vec3f center = ts.transform_point(ss.local_transform.position);
  const Vec3f lc = box_inv.transform_point(center);
  const Vec3f& half = bs.half_extents;
  const float radius = ss.radius;
  auto clampf = [](float v, float lo, float hi) { return std::max(lo, std::min(v, hi)); };
  const Vec3f closest(clampf(lc[0], -half[0], half[0]), clampf(lc[1], -half[1], half[1]),
                      clampf(lc[2], -half[2], half[2]));
  const Vec3f diff = lc - closest;
  const float dist_sq = diff.squaredNorm();
  if (dist_sq > radius * radius) return false;
  const Vec3f wclosest = box_world.transform_point(closest);
  const Vec3f wdiff = center - wclosest;
  const float dist = std::sqrt(dist_sq);
  Vec3f n;
  if (dist < 1e-6f) {
    const Vec3f fd(half[0] - std::fabs(lc[0]), half[1] - std::fabs(lc[1]), half[2] - std::fabs(lc[2]));
    int ax = 0;
    if (fd[1] < fd[ax]) ax = 1;
    if (fd[2] < fd[ax]) ax = 2;
    Vec3f ln;
    ln[ax] = lc[ax] > 0 ? 1.0f : -1.0f;
    n = box_world.transform_vector(ln);
  } else {
    n = wdiff / dist;
  }
  out.push_back(contact(wclosest, n, radius - dist, bs.body_index, ss.body_index,
                        combine_friction(bs.friction, ss.friction), combine_restitution(bs.restitution, ss.restitution)));
  return true;
}

// narrowphase.cpp:176-209: every box corner behind the plane
bool box_plane(const CollisionShape& bs, const Transform& tb, const CollisionShape& p,
               std::vector<ContactPoint>& out) {
  const Transform box_world = tb * bs.local_transform;
  const Vec3f& h = bs.half_extents;
  bool hit = false;
  for (int i = 0; i < 8; ++i) {
    const Vec3f corner((i & 1) ? h[0] : -h[0], (i & 2) ? h[1] : -h[1], (i & 4) ? h[2] : -h[2]);
    const Vec3f wc = box_world.transform_point(corner);
    const float distance = p.normal.dot(wc) - p.offset;
    if (distance < 0.0f) {
      out.push_back(contact(wc - p.normal * distance, p.normal, -distance, p.body_index, bs.body_index,
                            combine_friction(bs.friction, p.friction),
                            combine_restitution(bs.restitution, p.restitution)));
      hit = true;
    }
  }
  return hit;
}

// narrowphase.cpp:224-381: separating-axis test over the 15 axes (3 face axes
// of each box, 9 edge cross products; an edge axis must beat the best face
// axis by 1e-4 m to win), then contacts: for a face axis, up to 4 corners of
// the incident box inside the reference box (+0.01 m), each shifted half the
// penetration along the axis, or the centres' midpoint when none is; for an
// edge axis, the midpoint.
bool box_box(const CollisionShape& a, const Transform& ta, const CollisionShape& b, const Transform& tb,
             std::vector<ContactPoint>& out) {
  const Transform wa = ta * a.local_transform;
  const Transform wb = tb * b.local_transform;
  const Vec3f& ha = a.half_extents;
  const Vec3f& hb = b.half_extents;
  const Mat3f Ra = wa.rotation_matrix(), Rb = wb.rotation_matrix();
  const Vec3f pa = wa.position, pb = wb.position;
  const Vec3f d = pb - pa;
  const Vec3f axa[3] = {Ra.col(0), Ra.col(1), Ra.col(2)};
  const Vec3f axb[3] = {Rb.col(0), Rb.col(1), Rb.col(2)};
  const Mat3f R = Ra.transpose() * Rb;
  Mat3f absR;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) absR(i, j) = std::fabs(R(i, j)) + 1e-6f;
  const Vec3f t = Ra.transpose() * d;
  float min_pen = 3.40282347e38f;
  Vec3f best;
  int best_idx = -1;
  for (int i = 0; i < 3; ++i) {
    const float ra = ha[i];
    const float rb = hb[0] * absR(i, 0) + hb[1] * absR(i, 1) + hb[2] * absR(i, 2);
    const float pen = (ra + rb) - std::fabs(t[i]);
    if (pen < 0) return false;
    if (pen < min_pen) {
      min_pen = pen;
      best = t[i] < 0 ? -axa[i] : axa[i];
      best_idx = i;
    }
  }
  for (int i = 0; i < 3; ++i) {
    const float ra = ha[0] * absR(0, i) + ha[1] * absR(1, i) + ha[2] * absR(2, i);
    const float rb = hb[i];
    const float sp = R.col(i).dot(t);
    const float pen = (ra + rb) - std::fabs(sp);
    if (pen < 0) return false;
    if (pen < min_pen) {
      min_pen = pen;
      best = sp < 0 ? -axb[i] : axb[i];
      best_idx = 3 + i;
    }
  }
  const float edge_margin = 1e-4f;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      Vec3f axis = axa[i].cross(axb[j]);
      const float len = axis.norm();
      if (len < 1e-6f) continue;
      axis = axis / len;
      float ra = 0, rb = 0;
      for (int k = 0; k < 3; ++k) {
        ra += ha[k] * std::fabs(axis.dot(axa[k]));
        rb += hb[k] * std::fabs(axis.dot(axb[k]));
      }
      const float sp = axis.dot(d);
      const float pen = (ra + rb) - std::fabs(sp);
      if (pen < 0) return false;
      if (pen < min_pen - edge_margin) {
        min_pen = pen;
        best = sp < 0 ? -axis : axis;
        best_idx = 6 + i * 3 + j;
      }
    }
  const float fr = combine_friction(a.friction, b.friction), re = combine_restitution(a.restitution, b.restitution);
  const size_t before = out.size();
  if (best_idx < 6) {
    const bool a_ref = best_idx < 3;
    const Vec3f& ref_h = a_ref ? ha : hb;
    const Vec3f& inc_h = a_ref ? hb : ha;
    const Transform& ref_x = a_ref ? wa : wb;
    const Transform& inc_x = a_ref ? wb : wa;
    const float eps_in = 0.01f;
    const float cpen = std::max(0.0f, min_pen);
    int nface = 0;
    for (int i = 0; i < 8; ++i) {
      const Vec3f corner((i & 1) ? inc_h[0] : -inc_h[0], (i & 2) ? inc_h[1] : -inc_h[1],
                         (i & 4) ? inc_h[2] : -inc_h[2]);
      const Vec3f wc = inc_x.transform_point(corner);
      const Vec3f lc = ref_x.inverse().transform_point(wc);
      if (std::fabs(lc[0]) <= ref_h[0] + eps_in && std::fabs(lc[1]) <= ref_h[1] + eps_in &&
          std::fabs(lc[2]) <= ref_h[2] + eps_in) {
        const float off = 0.5f * cpen;
        out.push_back(contact(a_ref ? wc + best * off : wc - best * off, best, cpen, a.body_index, b.body_index, fr, re));
        if (++nface >= 4) break;
      }
    }
    if (out.size() == before) out.push_back(contact(pa + d * 0.5f, best, min_pen, a.body_index, b.body_index, fr, re));
  } else {
    out.push_back(contact(pa + d * 0.5f, best, min_pen, a.body_index, b.body_index, fr, re));
  }
  return out.size() > before;
}

}  // namespace

bool collide_shapes(const CollisionShape& a, const Transform& ta, const CollisionShape& b, const Transform& tb,
                    std::vector<ContactPoint>& contacts) {
  const ShapeType A = a.type, B = b.type;
  if (A == ShapeType::Sphere && B == ShapeType::Sphere) return sphere_sphere(a, ta, b, tb, contacts);
  if (A == ShapeType::Sphere && B == ShapeType::Plane) return sphere_plane(a, ta, b, contacts);
  if (A == ShapeType::Plane && B == ShapeType::Sphere) return sphere_plane(b, tb, a, contacts);
  if (A == ShapeType::Box && B == ShapeType::Sphere) return box_sphere(a, ta, b, tb, contacts);
  if (A == ShapeType::Sphere && B == ShapeType::Box) {
    // the box's contact, seen from the sphere: every contact in the list has
    // its normal negated and its bodies swapped (narrowphase.cpp:408-414)
    const bool r = box_sphere(b, tb, a, ta, contacts);
    for (auto& cp : contacts) {
      cp.normal = -cp.normal;
      std::swap(cp.body_a, cp.body_b);
    }
    return r;
  }
  if (A == ShapeType::Box && B == ShapeType::Plane) return box_plane(a, ta, b, contacts);
  if (A == ShapeType::Plane && B == ShapeType::Box) return box_plane(b, tb, a, contacts);
  if (A == ShapeType::Box && B == ShapeType::Box) return box_box(a, ta, b, tb, contacts);
  return false;
}

// ------------------------------------------------------------- solver
// FreeBodySolver::solve (free_body_solver.cpp:30-75, 89-210, 212-330):
// pre-step per contact (offsets, world inverse inertias, effective masses,
// tangent basis, Baumgarte + restitution bias), a warm start with the
// contacts' accumulated impulses, then velocity iterations of sequential
// impulses: normal (accumulated impulse clamped at 0), then two friction
// directions clamped to the friction cone, relative velocity recomputed
// before each.
void World::solve_contacts(float dt) {
  auto& C = contacts_;
  if (C.empty()) return;
  const auto& bodies = model_.bodies;
  auto& lv = state_.linear_velocities;
  auto& av = state_.angular_velocities;
  const auto& tf = state_.transforms;
  auto inv_mass = [&](int b) { return b >= 0 ? bodies[b].inv_mass() : 0.0f; };
  auto dynamic = [&](int b) { return b >= 0 && !bodies[b].is_static(); };
  auto vel = [&](int b, const Vec3f& r) { return b >= 0 ? lv[b] + av[b].cross(r) : Vec3f::Zero(); };
  auto apply = [&](const ContactPoint& cp, const ConstraintData& cd, const Vec3f& imp) {
    if (dynamic(cp.body_a)) {
      lv[cp.body_a] -= imp * bodies[cp.body_a].inv_mass();
      av[cp.body_a] -= cd.inv_I_a_world * cd.r_a.cross(imp);
    }
    if (dynamic(cp.body_b)) {
      lv[cp.body_b] += imp * bodies[cp.body_b].inv_mass();
      av[cp.body_b] += cd.inv_I_b_world * cd.r_b.cross(imp);
    }
  };
  {
    PhaseScope p(&monitor_, "world.solver.pre_step");
    constraint_data_.resize(C.size());
    for (size_t i = 0; i < C.size(); ++i) {  // sleeping bodies are not modelled (sleep_enabled is off)
      auto& cp = C[i];
      auto& cd = constraint_data_[i];
      cd.r_a = cp.body_a >= 0 ? cp.position - tf[cp.body_a].position : Vec3f::Zero();
      cd.r_b = cp.body_b >= 0 ? cp.position - tf[cp.body_b].position : Vec3f::Zero();
      const float ima = inv_mass(cp.body_a), imb = inv_mass(cp.body_b);
      auto world_inv_inertia = [&](int b) {
        if (!dynamic(b)) return Mat3f::Zero();
        const Mat3f R = tf[b].rotation_matrix();
        return R * bodies[b].inv_inertia() * R.transpose();
      };
      cd.inv_I_a_world = world_inv_inertia(cp.body_a);
      cd.inv_I_b_world = world_inv_inertia(cp.body_b);
      auto eff = [&](const Vec3f& dir) {
        const Vec3f ra = cd.r_a.cross(dir), rb = cd.r_b.cross(dir);
        float k = ima + imb + ra.dot(cd.inv_I_a_world * ra) + rb.dot(cd.inv_I_b_world * rb);
        if (k > 0) k = 1.0f / k;
        return k;
      };
      cd.effective_mass_n = eff(cp.normal);
      cd.tangent1 = (std::fabs(cp.normal.x()) < 0.9f ? cp.normal.cross(Vec3f(1, 0, 0)) : cp.normal.cross(Vec3f(0, 1, 0)))
                        .normalized();
      cd.tangent2 = cp.normal.cross(cd.tangent1);
      cd.effective_mass_t1 = eff(cd.tangent1);
      cd.effective_mass_t2 = eff(cd.tangent2);
      const float pen = std::max(0.0f, cp.penetration - settings_.slop);
      cd.bias = settings_.baumgarte * pen / dt;
      const float rel_vn = cp.normal.dot(vel(cp.body_b, cd.r_b) - vel(cp.body_a, cd.r_a));
      if (rel_vn < -1.0f) cd.bias += cp.restitution * rel_vn;
      if (!settings_.warm_starting) {
        cp.accumulated_normal_impulse = 0.0f;
        cp.accumulated_tangent_impulse_1 = 0.0f;
        cp.accumulated_tangent_impulse_2 = 0.0f;
      }
    }
  }
  if (settings_.warm_starting) {
    PhaseScope p(&monitor_, "world.solver.warm_start");
    for (size_t i = 0; i < C.size(); ++i) {
      const auto& cp = C[i];
      const auto& cd = constraint_data_[i];
      if (cd.effective_mass_n == 0.0f) continue;
      const Vec3f imp = cp.normal * cp.accumulated_normal_impulse + cd.tangent1 * cp.accumulated_tangent_impulse_1 +
                        cd.tangent2 * cp.accumulated_tangent_impulse_2;
      apply(cp, cd, imp);
    }
  }
  for (int it = 0; it < settings_.velocity_iterations; ++it) {
    PhaseScope p(&monitor_, "world.solver.iteration.solve_velocity");
    for (size_t i = 0; i < C.size(); ++i) {
      auto& cp = C[i];
      const auto& cd = constraint_data_[i];
      if (cd.effective_mass_n == 0.0f) continue;
      const float vn = (vel(cp.body_b, cd.r_b) - vel(cp.body_a, cd.r_a)).dot(cp.normal);
      const float ln = cd.effective_mass_n * (-vn + cd.bias);
      const float old_n = cp.accumulated_normal_impulse;
      cp.accumulated_normal_impulse = std::max(0.0f, old_n + ln);
      apply(cp, cd, cp.normal * (cp.accumulated_normal_impulse - old_n));
      const float max_f = cp.friction * cp.accumulated_normal_impulse;
      const float vt1 = (vel(cp.body_b, cd.r_b) - vel(cp.body_a, cd.r_a)).dot(cd.tangent1);
      const float lt1 = cd.effective_mass_t1 * (-vt1);
      const float old1 = cp.accumulated_tangent_impulse_1;
      cp.accumulated_tangent_impulse_1 = std::max(-max_f, std::min(max_f, old1 + lt1));
      apply(cp, cd, cd.tangent1 * (cp.accumulated_tangent_impulse_1 - old1));
      const float vt2 = (vel(cp.body_b, cd.r_b) - vel(cp.body_a, cd.r_a)).dot(cd.tangent2);
      const float lt2 = cd.effective_mass_t2 * (-vt2);
      const float old2 = cp.accumulated_tangent_impulse_2;
      cp.accumulated_tangent_impulse_2 = std::max(-max_f, std::min(max_f, old2 + lt2));
      apply(cp, cd, cd.tangent2 * (cp.accumulated_tangent_impulse_2 - old2));
    }
  }
}

// ------------------------------------------------------------- pipeline
// World::step_rigid_pipeline (world.cpp:30-182) with sleeping off (the
// reference's default): integrate velocities, shape AABBs, sweep and prune,
// narrowphase per candidate pair (contact friction / restitution combined from
// the two shapes), contact solve, integrate positions, clear forces, metrics.
void World::step_rigid_pipeline(float dt) {
  PhaseScope total(&monitor_, "world.total");
  const int n = model_.num_bodies();
  int dynamic_bodies = 0;
  {
    PhaseScope p(&monitor_, "world.integrate_velocity");
    for (int i = 0; i < n; ++i) {  // SymplecticEuler::integrate_velocity (integrator.cpp:13-24)
      const auto& b = model_.bodies[i];
      if (b.is_static()) continue;
      ++dynamic_bodies;
      Vec3f& v = state_.linear_velocities[i];
      Vec3f& w = state_.angular_velocities[i];
      v += (gravity_ + state_.forces[i] * b.inv_mass()) * dt;
      w += b.inv_inertia() * state_.torques[i] * dt;  // body-frame inverse inertia, as the reference
      w *= 0.999f;
    }
  }
  const int ns = model_.num_shapes();
  std::vector<AABB> boxes(ns);
  std::vector<bool> stat(ns);
  {
    PhaseScope p(&monitor_, "world.broadphase.build_aabbs");
    for (int i = 0; i < ns; ++i) {
      const auto& s = model_.shapes[i];
      if (s.body_index >= 0) {
        boxes[i] = shape_aabb(s, state_.transforms[s.body_index]);
        stat[i] = model_.bodies[s.body_index].is_static();
      } else {
        boxes[i] = shape_aabb(s, Transform::identity());
        stat[i] = true;
      }
    }
  }
  {
    PhaseScope p(&monitor_, "world.broadphase.sap");
    broadphase_.update(boxes, stat);
  }
  contacts_.clear();
  {
    PhaseScope p(&monitor_, "world.narrowphase.total");
    std::vector<ContactPoint> fresh;
    for (const auto& pr : broadphase_.get_pairs()) {
      const auto& sa = model_.shapes[pr.body_a];
      const auto& sb = model_.shapes[pr.body_b];
      const Transform ta = sa.body_index >= 0 ? state_.transforms[sa.body_index] : Transform::identity();
      const Transform tb = sb.body_index >= 0 ? state_.transforms[sb.body_index] : Transform::identity();
      fresh.clear();
      if (collide_shapes(sa, ta, sb, tb, fresh))
        for (auto& cp : fresh) {
          cp.friction = combine_friction(sa.friction, sb.friction);
          cp.restitution = combine_restitution(sa.restitution, sb.restitution);
          contacts_.push_back(cp);
        }
    }
  }
  {
    PhaseScope p(&monitor_, "world.solver.total");
    solve_contacts(dt);
  }
  {
    PhaseScope p(&monitor_, "world.integrate_position");
    for (int i = 0; i < n; ++i) {  // SymplecticEuler::integrate_position (integrator.cpp:26-60)
      if (model_.bodies[i].is_static()) continue;
      Transform& t = state_.transforms[i];
      const Vec3f& w = state_.angular_velocities[i];
      t.position += state_.linear_velocities[i] * dt;
      Quatf& q = t.rotation;
      const float dw = 0.5f * (-w[0] * q.x_ - w[1] * q.y_ - w[2] * q.z_);
      const float dx = 0.5f * (w[0] * q.w_ + w[2] * q.y_ - w[1] * q.z_);
      const float dy = 0.5f * (w[1] * q.w_ - w[2] * q.x_ + w[0] * q.z_);
      const float dz = 0.5f * (w[2] * q.w_ + w[1] * q.x_ - w[0] * q.y_);
      q = Quatf(q.w_ + dw * dt, q.x_ + dx * dt, q.y_ + dy * dt, q.z_ + dz * dt).normalized();
    }
  }
  {
    PhaseScope p(&monitor_, "world.clear_forces");
    state_.clear_forces();
  }
  // World::record_world_metrics (world.cpp:218-229)
  monitor_.record_metric("bodies", n);
  monitor_.record_metric("dynamic_bodies", dynamic_bodies);
  monitor_.record_metric("shapes", ns);
  monitor_.record_metric("candidate_pairs", (double)broadphase_.get_pairs().size());
  monitor_.record_metric("contacts", (double)contacts_.size());
  monitor_.record_metric("solver_iterations", settings_.velocity_iterations);
}

}  // namespace novaphy
