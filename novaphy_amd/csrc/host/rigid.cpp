// rigid.cpp -- math helpers, model factories and the host rigid-body pipeline.
//
// The rigid pipeline stays on the host: it follows the stages of
// World::step_rigid_pipeline (src/sim/world.cpp:30-182) -- semi-implicit Euler
// velocity update with 0.999 angular damping (src/dynamics/integrator.cpp),
// contact generation for sphere / box / plane pairs
// (src/collision/narrowphase.cpp), a sequential-impulse solve with Baumgarte
// bias, restitution and two friction directions
// (src/dynamics/free_body_solver.cpp), position / quaternion integration and
// force clearing.  Broadphase is a plain AABB overlap test over shape pairs
// (a handful of bodies per fluid scene); box-box contacts use corner-in-box
// tests in both directions rather than the reference's 15-axis SAT.
#include <algorithm>

#include "npx_host.hpp"

namespace novaphy {

Mat3f Mat3f::inverse() const {
  const float a = m[0][0], b = m[0][1], c = m[0][2], d = m[1][0], e = m[1][1], f = m[1][2], g = m[2][0],
              h = m[2][1], i = m[2][2];
  const float A = e * i - f * h, B = -(d * i - f * g), Cc = d * h - e * g;
  const float det = a * A + b * B + c * Cc;
  Mat3f r;
  if (det == 0.0f) return r;
  const float inv = 1.0f / det;
  r.m[0][0] = A * inv;
  r.m[0][1] = -(b * i - c * h) * inv;
  r.m[0][2] = (b * f - c * e) * inv;
  r.m[1][0] = B * inv;
  r.m[1][1] = (a * i - c * g) * inv;
  r.m[1][2] = -(a * f - c * d) * inv;
  r.m[2][0] = Cc * inv;
  r.m[2][1] = -(a * h - b * g) * inv;
  r.m[2][2] = (a * e - b * d) * inv;
  return r;
}

Mat3f Quatf::matrix() const {
  Mat3f r;
  const float w = w_, x = x_, y = y_, z = z_;
  r.m[0][0] = 1 - 2 * (y * y + z * z);
  r.m[0][1] = 2 * (x * y - z * w);
  r.m[0][2] = 2 * (x * z + y * w);
  r.m[1][0] = 2 * (x * y + z * w);
  r.m[1][1] = 1 - 2 * (x * x + z * z);
  r.m[1][2] = 2 * (y * z - x * w);
  r.m[2][0] = 2 * (x * z - y * w);
  r.m[2][1] = 2 * (y * z + x * w);
  r.m[2][2] = 1 - 2 * (x * x + y * y);
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
static void check_body(const SimState& s, int i) {
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
void SimState::apply_force(int i, const Vec3f& f) {
  check_body(*this, i);
  forces[i] += f;
}
void SimState::apply_torque(int i, const Vec3f& t) {
  check_body(*this, i);
  torques[i] += t;
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

namespace {

struct Aabb {
  Vec3f lo, hi;
};

Transform shape_world(const CollisionShape& s, const SimState& st) {
  const Transform body = s.body_index >= 0 ? st.transforms[s.body_index] : Transform::identity();
  return body * s.local_transform;
}

Aabb shape_aabb(const CollisionShape& s, const Transform& w) {
  Aabb a;
  if (s.type == ShapeType::Plane) {
    a.lo = Vec3f(-1e6f, -1e6f, -1e6f);
    a.hi = Vec3f(1e6f, 1e6f, 1e6f);
  } else if (s.type == ShapeType::Sphere) {
    const Vec3f r(s.radius, s.radius, s.radius);
    a.lo = w.position - r;
    a.hi = w.position + r;
  } else {
    const Mat3f R = w.rotation_matrix();
    Vec3f e;
    for (int i = 0; i < 3; ++i)
      e[i] = std::fabs(R(i, 0)) * s.half_extents[0] + std::fabs(R(i, 1)) * s.half_extents[1] +
             std::fabs(R(i, 2)) * s.half_extents[2];
    a.lo = w.position - e;
    a.hi = w.position + e;
  }
  return a;
}

bool overlap(const Aabb& a, const Aabb& b) {
  for (int i = 0; i < 3; ++i)
    if (a.hi[i] < b.lo[i] || b.hi[i] < a.lo[i]) return false;
  return true;
}

ContactPoint make_cp(const Vec3f& pos, const Vec3f& n, float pen, int ba, int bb, const CollisionShape& sa,
                     const CollisionShape& sb) {
  ContactPoint c;
  c.position = pos;
  c.normal = n;
  c.penetration = pen;
  c.body_a = ba;
  c.body_b = bb;
  c.friction = std::sqrt(sa.friction * sb.friction);     // combine_friction
  c.restitution = std::max(sa.restitution, sb.restitution);  // combine_restitution
  return c;
}

// normal points from A toward B (contact.h)
void sphere_sphere(const CollisionShape& a, const Transform& ta, const CollisionShape& b, const Transform& tb,
                   std::vector<ContactPoint>& out) {
  const Vec3f d = tb.position - ta.position;
  const float dist = d.norm(), rs = a.radius + b.radius;
  if (dist >= rs) return;
  const Vec3f n = dist > 1e-6f ? d * (1.0f / dist) : Vec3f(0, 1, 0);
  const float pen = rs - dist;
  out.push_back(make_cp(ta.position + n * (a.radius - 0.5f * pen), n, pen, a.body_index, b.body_index, a, b));
}

void sphere_plane(const CollisionShape& s, const Transform& ts, const CollisionShape& p,
                  std::vector<ContactPoint>& out) {
  const float dist = p.normal.dot(ts.position) - p.offset;
  if (dist > s.radius) return;
  out.push_back(make_cp(ts.position - p.normal * dist, p.normal, s.radius - dist, p.body_index, s.body_index, p, s));
}

void box_plane(const CollisionShape& b, const Transform& tb, const CollisionShape& p,
               std::vector<ContactPoint>& out) {
  const Vec3f& he = b.half_extents;
  for (int i = 0; i < 8; ++i) {
    const Vec3f c((i & 1) ? he[0] : -he[0], (i & 2) ? he[1] : -he[1], (i & 4) ? he[2] : -he[2]);
    const Vec3f w = tb.transform_point(c);
    const float d = p.normal.dot(w) - p.offset;
    if (d < 0.0f) out.push_back(make_cp(w - p.normal * d, p.normal, -d, p.body_index, b.body_index, p, b));
  }
}

// sphere against oriented box: closest point in the box frame
void box_sphere(const CollisionShape& b, const Transform& tb, const CollisionShape& s, const Transform& ts,
                std::vector<ContactPoint>& out) {
  const Transform inv = tb.inverse();
  const Vec3f lc = inv.transform_point(ts.position);
  const Vec3f& he = b.half_extents;
  Vec3f cl;
  bool inside = true;
  for (int i = 0; i < 3; ++i) {
    cl[i] = std::min(std::max(lc[i], -he[i]), he[i]);
    if (cl[i] != lc[i]) inside = false;
  }
  Vec3f n;
  float pen;
  Vec3f wcl = tb.transform_point(cl);
  if (inside) {
    int ax = 0;
    float best = 1e30f;
    for (int i = 0; i < 3; ++i) {
      const float d = he[i] - std::fabs(lc[i]);
      if (d < best) best = d, ax = i;
    }
    Vec3f ln;
    ln[ax] = lc[ax] > 0 ? 1.0f : -1.0f;
    n = tb.transform_vector(ln);
    pen = s.radius + best;
  } else {
    const Vec3f d = ts.position - wcl;
    const float dist = d.norm();
    if (dist >= s.radius) return;
    n = d * (1.0f / std::max(dist, 1e-9f));
    pen = s.radius - dist;
  }
  out.push_back(make_cp(wcl, n, pen, b.body_index, s.body_index, b, s));
}

// corners of box A inside box B (normal along B's least-penetrated face)
void box_corners_in_box(const CollisionShape& a, const Transform& ta, const CollisionShape& b,
                        const Transform& tb, bool flip, std::vector<ContactPoint>& out) {
  const Transform inv = tb.inverse();
  const Vec3f& ha = a.half_extents;
  const Vec3f& hb = b.half_extents;
  for (int i = 0; i < 8; ++i) {
    const Vec3f c((i & 1) ? ha[0] : -ha[0], (i & 2) ? ha[1] : -ha[1], (i & 4) ? ha[2] : -ha[2]);
    const Vec3f w = ta.transform_point(c);
    const Vec3f l = inv.transform_point(w);
    int ax = -1;
    float best = 1e30f;
    for (int k = 0; k < 3; ++k) {
      const float d = hb[k] - std::fabs(l[k]);
      if (d <= 0.0f) {
        ax = -2;
        break;
      }
      if (d < best) best = d, ax = k;
    }
    if (ax < 0) continue;
    Vec3f ln;
    ln[ax] = l[ax] > 0 ? 1.0f : -1.0f;
    Vec3f n = tb.transform_vector(ln);  // outward from B, i.e. from B toward A's corner
    if (!flip) {
      // contact normal from A toward B
      out.push_back(make_cp(w, -n, best, a.body_index, b.body_index, a, b));
    } else {
      out.push_back(make_cp(w, n, best, b.body_index, a.body_index, b, a));
    }
  }
}

}  // namespace

void World::collide(std::vector<ContactPoint>& out) {
  const int ns = model_.num_shapes();
  std::vector<Transform> w(ns);
  std::vector<Aabb> box(ns);
  std::vector<bool> stat(ns);
  for (int i = 0; i < ns; ++i) {
    const auto& s = model_.shapes[i];
    w[i] = shape_world(s, state_);
    box[i] = shape_aabb(s, w[i]);
    stat[i] = s.body_index < 0 || model_.bodies[s.body_index].is_static();
  }
  for (int i = 0; i < ns; ++i)
    for (int j = i + 1; j < ns; ++j) {
      if (stat[i] && stat[j]) continue;
      const auto& a = model_.shapes[i];
      const auto& b = model_.shapes[j];
      if (a.body_index >= 0 && a.body_index == b.body_index) continue;
      if (!overlap(box[i], box[j])) continue;
      const size_t before = out.size();
      const ShapeType A = a.type, B = b.type;
      if (A == ShapeType::Sphere && B == ShapeType::Sphere)
        sphere_sphere(a, w[i], b, w[j], out);
      else if (A == ShapeType::Sphere && B == ShapeType::Plane)
        sphere_plane(a, w[i], b, out);
      else if (A == ShapeType::Plane && B == ShapeType::Sphere)
        sphere_plane(b, w[j], a, out);
      else if (A == ShapeType::Box && B == ShapeType::Sphere)
        box_sphere(a, w[i], b, w[j], out);
      else if (A == ShapeType::Sphere && B == ShapeType::Box)
        box_sphere(b, w[j], a, w[i], out);
      else if (A == ShapeType::Box && B == ShapeType::Plane)
        box_plane(a, w[i], b, out);
      else if (A == ShapeType::Plane && B == ShapeType::Box)
        box_plane(b, w[j], a, out);
      else if (A == ShapeType::Box && B == ShapeType::Box) {
        box_corners_in_box(a, w[i], b, w[j], false, out);
        box_corners_in_box(b, w[j], a, w[i], true, out);
      }
      for (size_t k = before; k < out.size(); ++k) {
        out[k].shape_a = (out[k].body_a == a.body_index) ? i : j;
        out[k].shape_b = (out[k].shape_a == i) ? j : i;
      }
    }
}

void World::solve_contacts(float dt) {
  if (contacts_.empty()) return;
  struct CD {
    Vec3f ra, rb, t1, t2;
    Mat3f Ia, Ib;
    float mn = 0, mt1 = 0, mt2 = 0, bias = 0;
  };
  std::vector<CD> cd(contacts_.size());
  auto inv_mass = [&](int b) { return b >= 0 ? model_.bodies[b].inv_mass() : 0.0f; };
  auto dyn = [&](int b) { return b >= 0 && !model_.bodies[b].is_static(); };
  auto vel_at = [&](int b, const Vec3f& r) {
    return b >= 0 ? state_.linear_velocities[b] + state_.angular_velocities[b].cross(r) : Vec3f::Zero();
  };
  auto inv_world_inertia = [&](int b) {
    if (!dyn(b)) return Mat3f::Zero();
    const Mat3f R = state_.transforms[b].rotation_matrix();
    return R * model_.bodies[b].inv_inertia() * R.transpose();
  };
  auto eff = [&](const CD& d, const ContactPoint& c, const Vec3f& dir) {
    const Vec3f a = d.ra.cross(dir), b = d.rb.cross(dir);
    float k = inv_mass(c.body_a) + inv_mass(c.body_b) + a.dot(d.Ia * a) + b.dot(d.Ib * b);
    return k > 0 ? 1.0f / k : 0.0f;
  };
  for (size_t i = 0; i < contacts_.size(); ++i) {  // pre_step (free_body_solver.cpp:89-210)
    auto& c = contacts_[i];
    auto& d = cd[i];
    d.ra = c.body_a >= 0 ? c.position - state_.transforms[c.body_a].position : Vec3f::Zero();
    d.rb = c.body_b >= 0 ? c.position - state_.transforms[c.body_b].position : Vec3f::Zero();
    d.Ia = inv_world_inertia(c.body_a);
    d.Ib = inv_world_inertia(c.body_b);
    d.mn = eff(d, c, c.normal);
    d.t1 = (std::fabs(c.normal[0]) < 0.9f ? c.normal.cross(Vec3f(1, 0, 0)) : c.normal.cross(Vec3f(0, 1, 0)))
               .normalized();
    d.t2 = c.normal.cross(d.t1);
    d.mt1 = eff(d, c, d.t1);
    d.mt2 = eff(d, c, d.t2);
    d.bias = settings_.baumgarte * std::max(0.0f, c.penetration - settings_.slop) / dt;
    const float vn = c.normal.dot(vel_at(c.body_b, d.rb) - vel_at(c.body_a, d.ra));
    if (vn < -1.0f) d.bias += c.restitution * vn;
  }
  auto apply = [&](const ContactPoint& c, const CD& d, const Vec3f& imp) {
    if (dyn(c.body_a)) {
      state_.linear_velocities[c.body_a] -= imp * inv_mass(c.body_a);
      state_.angular_velocities[c.body_a] -= d.Ia * d.ra.cross(imp);
    }
    if (dyn(c.body_b)) {
      state_.linear_velocities[c.body_b] += imp * inv_mass(c.body_b);
      state_.angular_velocities[c.body_b] += d.Ib * d.rb.cross(imp);
    }
  };
  for (int it = 0; it < settings_.velocity_iterations; ++it) {  // solve_velocity (:212-330)
    for (size_t i = 0; i < contacts_.size(); ++i) {
      auto& c = contacts_[i];
      auto& d = cd[i];
      if (d.mn == 0.0f) continue;
      float vn = (vel_at(c.body_b, d.rb) - vel_at(c.body_a, d.ra)).dot(c.normal);
      float ln = d.mn * (-vn + d.bias);
      const float old = c.acc_n;
      c.acc_n = std::max(0.0f, old + ln);
      apply(c, d, c.normal * (c.acc_n - old));
      const float maxf = c.friction * c.acc_n;
      for (int k = 0; k < 2; ++k) {
        const Vec3f& t = k == 0 ? d.t1 : d.t2;
        float& acc = k == 0 ? c.acc_t1 : c.acc_t2;
        const float m = k == 0 ? d.mt1 : d.mt2;
        const float vt = (vel_at(c.body_b, d.rb) - vel_at(c.body_a, d.ra)).dot(t);
        const float o = acc;
        acc = std::max(-maxf, std::min(maxf, o + m * (-vt)));
        apply(c, d, t * (acc - o));
      }
    }
  }
}

void World::step_rigid_pipeline(float dt) {
  PhaseScope total(&monitor_, "world.total");
  const int n = model_.num_bodies();
  {
    PhaseScope p(&monitor_, "world.integrate_velocity");
    for (int i = 0; i < n; ++i) {
      const auto& b = model_.bodies[i];
      if (b.is_static()) continue;
      state_.linear_velocities[i] += (gravity_ + state_.forces[i] * b.inv_mass()) * dt;
      const Mat3f R = state_.transforms[i].rotation_matrix();
      const Mat3f Iw = R * b.inv_inertia() * R.transpose();
      state_.angular_velocities[i] += Iw * state_.torques[i] * dt;
      state_.angular_velocities[i] *= 0.999f;
    }
  }
  contacts_.clear();
  {
    PhaseScope p(&monitor_, "world.broadphase.sap");
    collide(contacts_);
  }
  {
    PhaseScope p(&monitor_, "world.solver.total");
    solve_contacts(dt);
  }
  {
    PhaseScope p(&monitor_, "world.integrate_position");
    for (int i = 0; i < n; ++i) {
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
  monitor_.record_metric("bodies", n);
  monitor_.record_metric("contacts", (double)contacts_.size());
}

}  // namespace novaphy
