// rigid_oracle.cpp -- CPU restatement of NovaPhy's host rigid-body step.
//
// TEST INFRASTRUCTURE ONLY, like pbf_oracle.cpp: the checker of the product's
// host rigid pipeline (novaphy_amd/csrc/host/rigid.cpp) and of coupled
// FluidWorld rollouts.  Nothing in novaphy_amd/ links or imports it.
//
// It restates, FP32 and in the reference's operation order:
//   src/sim/world.cpp:30-182                 World::step_rigid_pipeline (sleep off)
//   src/dynamics/integrator.cpp:13-60        SymplecticEuler
//   src/collision/broadphase.cpp:16-118      SweepAndPrune::update
//   src/collision/narrowphase.cpp:25-432     sphere/box/plane pairs, 15-axis SAT
//   src/dynamics/free_body_solver.cpp:30-330 FreeBodySolver (warm start, PGS)
//   include/novaphy/core/shape.h:62-75       CollisionShape::compute_aabb
//   include/novaphy/core/aabb.h:90-140       AABB::from_sphere / from_oriented_box
//   include/novaphy/core/body.h:25-35        inv_mass / inv_inertia / is_static
//   include/novaphy/math/math_types.h:38-90  Transform
//   include/novaphy/core/contact.h:55-61     combine_friction / restitution
// Eigen is not available; its evaluation order is restated as follows (the
// reference binary is not built here, so these last-bit choices are unpinned
// against it and pinned instead by the closed forms and reference test
// properties in tests/test_rigid_oracle_pins.py): 3-term dot products and
// matrix products sum a0 + (a1 + a2); toRotationMatrix uses the doubled
// components (1 - (tyy + tzz), txy - twz, ...); the 3x3 inverse is the
// cofactor form with one reciprocal of the determinant; quaternion products
// follow the textbook Hamilton expansion left to right; normalisation divides
// each component by sqrt(w^2 + (x^2 + (y^2 + z^2))).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

extern "C" {

typedef struct orc_rbody {
  float mass;
  float inertia[9];  // body-frame inertia, row-major
} orc_rbody;

typedef struct orc_rshape {
  int32_t type;  // 0 box, 1 sphere, 2 plane (ShapeType order, shape.h:12)
  int32_t body;  // -1: world-owned
  float half[3];
  float radius;
  float normal[3];
  float offset;
  float local_pos[3];
  float local_rot[4];  // w, x, y, z
  float friction;
  float restitution;
} orc_rshape;

typedef struct orc_rsettings {
  int32_t velocity_iterations;
  float baumgarte;
  float slop;
  int32_t warm_starting;
} orc_rsettings;

typedef struct orc_contact {
  float position[3];
  float normal[3];
  float penetration;
  int32_t body_a, body_b;
  float friction, restitution;
  float acc_n, acc_t1, acc_t2;
} orc_contact;

}  // extern "C"

namespace orr {

struct V {
  float x, y, z;
};
inline V operator+(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V operator-(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V operator-(V a) { return {-a.x, -a.y, -a.z}; }
inline V operator*(V a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V operator/(V a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline float dot(V a, V b) { return a.x * b.x + (a.y * b.y + a.z * b.z); }
inline V cross(V a, V b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float at(V a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
inline void set(V& a, int i, float v) { (i == 0 ? a.x : (i == 1 ? a.y : a.z)) = v; }
inline V normalized(V a) {
  const float z = dot(a, a);
  if (z > 0.0f) {
    const float l = std::sqrt(z);
    return {a.x / l, a.y / l, a.z / l};
  }
  return a;
}

struct M {
  float a[3][3];
};
inline M zero() { return M{{{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}}; }
inline V mul(const M& m, V v) {
  return {m.a[0][0] * v.x + (m.a[0][1] * v.y + m.a[0][2] * v.z), m.a[1][0] * v.x + (m.a[1][1] * v.y + m.a[1][2] * v.z),
          m.a[2][0] * v.x + (m.a[2][1] * v.y + m.a[2][2] * v.z)};
}
inline M mul(const M& p, const M& q) {
  M r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i][j] = p.a[i][0] * q.a[0][j] + (p.a[i][1] * q.a[1][j] + p.a[i][2] * q.a[2][j]);
  return r;
}
inline M transpose(const M& m) {
  M r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i][j] = m.a[j][i];
  return r;
}
inline V col(const M& m, int j) { return {m.a[0][j], m.a[1][j], m.a[2][j]}; }
M inverse(const M& m) {
  auto cof = [&](int i, int j) {
    const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    return m.a[i1][j1] * m.a[i2][j2] - m.a[i1][j2] * m.a[i2][j1];
  };
  const float c00 = cof(0, 0), c10 = cof(1, 0), c20 = cof(2, 0);
  const float det = c00 * m.a[0][0] + (c10 * m.a[1][0] + c20 * m.a[2][0]);
  const float inv = 1.0f / det;
  M r;
  r.a[0][0] = c00 * inv;
  r.a[0][1] = c10 * inv;
  r.a[0][2] = c20 * inv;
  for (int j = 1; j < 3; ++j)
    for (int i = 0; i < 3; ++i) r.a[j][i] = cof(i, j) * inv;
  return r;
}

struct Q {
  float w, x, y, z;
};
inline Q qnormalized(Q q) {
  const float n = std::sqrt(q.w * q.w + (q.x * q.x + (q.y * q.y + q.z * q.z)));
  return {q.w / n, q.x / n, q.y / n, q.z / n};
}
inline Q qmul(Q a, Q b) {
  return {a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z, a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
          a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z, a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x};
}
inline V qrot(Q q, V v) {  // Eigen's _transformVector
  const V u{q.x, q.y, q.z};
  V uv = cross(u, v);
  uv = uv + uv;
  return (v + uv * q.w) + cross(u, uv);
}
M qmatrix(Q q) {
  const float tx = 2.0f * q.x, ty = 2.0f * q.y, tz = 2.0f * q.z;
  const float twx = tx * q.w, twy = ty * q.w, twz = tz * q.w;
  const float txx = tx * q.x, txy = ty * q.x, txz = tz * q.x;
  const float tyy = ty * q.y, tyz = tz * q.y, tzz = tz * q.z;
  return M{{{1.0f - (tyy + tzz), txy - twz, txz + twy},
            {txy + twz, 1.0f - (txx + tzz), tyz - twx},
            {txz - twy, tyz + twx, 1.0f - (txx + tyy)}}};
}

// Transform (math_types.h:38-90): the (position, rotation) constructor
// normalises the rotation.
struct T {
  V p{0, 0, 0};
  Q q{1, 0, 0, 0};
};
inline T make_t(V p, Q q) { return T{p, qnormalized(q)}; }
inline T tmul(const T& a, const T& b) { return make_t(a.p + qrot(a.q, b.p), qmul(a.q, b.q)); }
inline T tinv(const T& a) {
  const Q c{a.q.w, -a.q.x, -a.q.y, -a.q.z};
  return make_t(qrot(c, -a.p), c);
}
inline V tpoint(const T& a, V v) { return a.p + qrot(a.q, v); }

struct Box {
  V lo{3.40282347e38f, 3.40282347e38f, 3.40282347e38f}, hi{-3.40282347e38f, -3.40282347e38f, -3.40282347e38f};
};

struct Shape {
  int type, body;
  V half, normal;
  float radius, offset, friction, restitution;
  T local;
};

struct Contact {
  V pos{0, 0, 0}, n{0, 0, 0};
  float pen = 0.0f;
  int a = -1, b = -1;
  float fr = 0.5f, re = 0.0f;
  float an = 0.0f, at1 = 0.0f, at2 = 0.0f;
};

Box shape_box(const Shape& s, const T& body) {  // shape.h:62-75
  const T w = tmul(body, s.local);
  Box r;
  if (s.type == 0) {
    for (int i = 0; i < 8; ++i) {
      const V c{(i & 1) ? s.half.x : -s.half.x, (i & 2) ? s.half.y : -s.half.y, (i & 4) ? s.half.z : -s.half.z};
      const V wc = tpoint(w, c);
      r.lo = {std::min(r.lo.x, wc.x), std::min(r.lo.y, wc.y), std::min(r.lo.z, wc.z)};
      r.hi = {std::max(r.hi.x, wc.x), std::max(r.hi.y, wc.y), std::max(r.hi.z, wc.z)};
    }
  } else if (s.type == 1) {
    const V e{s.radius, s.radius, s.radius};
    r.lo = w.p - e;
    r.hi = w.p + e;
  } else {
    r.lo = {-1e6f, -1e6f, -1e6f};
    r.hi = {1e6f, 1e6f, 1e6f};
  }
  return r;
}

// broadphase.cpp:16-118 (the sort method does not change the sorted,
// deduplicated pair list, so one std::sort per frame stands for both paths)
std::vector<std::pair<int, int>> sap(const std::vector<Box>& bx, const std::vector<char>& st) {
  struct E {
    float v;
    int i;
    bool mn;
  };
  const int n = (int)bx.size();
  std::vector<E> ex(2 * (size_t)n);
  for (int i = 0; i < n; ++i) {
    ex[2 * i] = {bx[i].lo.x, i, true};
    ex[2 * i + 1] = {bx[i].hi.x, i, false};
  }
  std::sort(ex.begin(), ex.end(), [](const E& a, const E& b) { return a.v < b.v; });
  std::vector<std::pair<int, int>> pr;
  for (size_t i = 0; i < ex.size(); ++i) {
    if (!ex[i].mn) continue;
    const int a = ex[i].i;
    const float mx = bx[a].hi.x;
    for (size_t j = i + 1; j < ex.size(); ++j) {
      if (ex[j].v > mx) break;
      if (!ex[j].mn) continue;
      const int b = ex[j].i;
      if (a == b) continue;
      if (st[a] && st[b]) continue;
      if (bx[a].lo.y <= bx[b].hi.y && bx[a].hi.y >= bx[b].lo.y && bx[a].lo.z <= bx[b].hi.z && bx[a].hi.z >= bx[b].lo.z)
        pr.push_back({std::min(a, b), std::max(a, b)});
    }
  }
  std::sort(pr.begin(), pr.end());
  pr.erase(std::unique(pr.begin(), pr.end()), pr.end());
  return pr;
}

inline float cfr(float a, float b) { return std::sqrt(a * b); }
inline float cre(float a, float b) { return std::max(a, b); }

Contact mk(V pos, V n, float pen, int a, int b, float fr, float re) {
  Contact c;
  c.pos = pos;
  c.n = n;
  c.pen = pen;
  c.a = a;
  c.b = b;
  c.fr = fr;
  c.re = re;
  return c;
}

bool c_sphere_sphere(const Shape& a, const T& ta, const Shape& b, const T& tb, std::vector<Contact>& out) {
  const V ca = tpoint(ta, a.local.p), cb = tpoint(tb, b.local.p);
  const V diff = cb - ca;
  const float dsq = dot(diff, diff), rs = a.radius + b.radius;
  if (dsq > rs * rs) return false;
  const float dist = std::sqrt(dsq);
  const V n = dist < 1e-6f ? V{0, 1, 0} : diff / dist;
  const float pen = rs - dist;
  out.push_back(mk(ca + n * (a.radius - pen * 0.5f), n, pen, a.body, b.body, cfr(a.friction, b.friction),
                   cre(a.restitution, b.restitution)));
  return true;
}

bool c_sphere_plane(const Shape& s, const T& ts, const Shape& p, std::vector<Contact>& out) {
  const V c = tpoint(ts, s.local.p);
  const float d = dot(p.normal, c) - p.offset;
  if (d > s.radius) return false;
  out.push_back(mk(c - p.normal * d, p.normal, s.radius - d, p.body, s.body, cfr(s.friction, p.friction),
                   cre(s.restitution, p.restitution)));
  return true;
}

bool c_box_sphere(const Shape& bs, const T& tb, const Shape& ss, const T& ts, std::vector<Contact>& out) {
  const T bw = tmul(tb, bs.local);
  const T bi = tinv(bw);
  const V c = tpoint(ts, ss.local.p);
  const V lc = tpoint(bi, c);
  const V h = bs.half;
  auto cl = [](float v, float lo, float hi) { return std::max(lo, std::min(v, hi)); };
  const V cp{cl(lc.x, -h.x, h.x), cl(lc.y, -h.y, h.y), cl(lc.z, -h.z, h.z)};
  const V diff = lc - cp;
  const float dsq = dot(diff, diff);
  if (dsq > ss.radius * ss.radius) return false;
  const V wcp = tpoint(bw, cp);
  const V wd = c - wcp;
  const float dist = std::sqrt(dsq);
  V n;
  if (dist < 1e-6f) {
    const V fd{h.x - std::fabs(lc.x), h.y - std::fabs(lc.y), h.z - std::fabs(lc.z)};
    int ax = 0;
    if (fd.y < at(fd, ax)) ax = 1;
    if (fd.z < at(fd, ax)) ax = 2;
    V ln{0, 0, 0};
    set(ln, ax, at(lc, ax) > 0 ? 1.0f : -1.0f);
    n = qrot(bw.q, ln);
  } else {
    n = wd / dist;
  }
  out.push_back(mk(wcp, n, ss.radius - dist, bs.body, ss.body, cfr(bs.friction, ss.friction),
                   cre(bs.restitution, ss.restitution)));
  return true;
}

bool c_box_plane(const Shape& bs, const T& tb, const Shape& p, std::vector<Contact>& out) {
  const T bw = tmul(tb, bs.local);
  const V h = bs.half;
  bool any = false;
  for (int i = 0; i < 8; ++i) {
    const V c{(i & 1) ? h.x : -h.x, (i & 2) ? h.y : -h.y, (i & 4) ? h.z : -h.z};
    const V wc = tpoint(bw, c);
    const float d = dot(p.normal, wc) - p.offset;
    if (d < 0.0f) {
      out.push_back(mk(wc - p.normal * d, p.normal, -d, p.body, bs.body, cfr(bs.friction, p.friction),
                       cre(bs.restitution, p.restitution)));
      any = true;
    }
  }
  return any;
}

bool c_box_box(const Shape& a, const T& ta, const Shape& b, const T& tb, std::vector<Contact>& out) {
  const T wa = tmul(ta, a.local), wb = tmul(tb, b.local);
  const V ha = a.half, hb = b.half;
  const M Ra = qmatrix(wa.q), Rb = qmatrix(wb.q);
  const V d = wb.p - wa.p;
  const V axa[3] = {col(Ra, 0), col(Ra, 1), col(Ra, 2)};
  const V axb[3] = {col(Rb, 0), col(Rb, 1), col(Rb, 2)};
  const M R = mul(transpose(Ra), Rb);
  M aR;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) aR.a[i][j] = std::fabs(R.a[i][j]) + 1e-6f;
  const V t = mul(transpose(Ra), d);
  float mp = 3.40282347e38f;
  V best{0, 0, 0};
  int bi = -1;
  for (int i = 0; i < 3; ++i) {  // face axes of A
    const float rb = hb.x * aR.a[i][0] + hb.y * aR.a[i][1] + hb.z * aR.a[i][2];
    const float pen = (at(ha, i) + rb) - std::fabs(at(t, i));
    if (pen < 0) return false;
    if (pen < mp) {
      mp = pen;
      best = at(t, i) < 0 ? -axa[i] : axa[i];
      bi = i;
    }
  }
  for (int i = 0; i < 3; ++i) {  // face axes of B
    const float ra = ha.x * aR.a[0][i] + ha.y * aR.a[1][i] + ha.z * aR.a[2][i];
    const float s = dot(col(R, i), t);
    const float pen = (ra + at(hb, i)) - std::fabs(s);
    if (pen < 0) return false;
    if (pen < mp) {
      mp = pen;
      best = s < 0 ? -axb[i] : axb[i];
      bi = 3 + i;
    }
  }
  for (int i = 0; i < 3; ++i)  // edge-edge axes
    for (int j = 0; j < 3; ++j) {
      V ax = cross(axa[i], axb[j]);
      const float len = std::sqrt(dot(ax, ax));
      if (len < 1e-6f) continue;
      ax = ax / len;
      float ra = 0, rb = 0;
      for (int k = 0; k < 3; ++k) {
        ra += at(ha, k) * std::fabs(dot(ax, axa[k]));
        rb += at(hb, k) * std::fabs(dot(ax, axb[k]));
      }
      const float s = dot(ax, d);
      const float pen = (ra + rb) - std::fabs(s);
      if (pen < 0) return false;
      if (pen < mp - 1e-4f) {
        mp = pen;
        best = s < 0 ? -ax : ax;
        bi = 6 + i * 3 + j;
      }
    }
  const float fr = cfr(a.friction, b.friction), re = cre(a.restitution, b.restitution);
  const size_t n0 = out.size();
  if (bi < 6) {
    const bool aref = bi < 3;
    const V rh = aref ? ha : hb, ih = aref ? hb : ha;
    const T& rx = aref ? wa : wb;
    const T& ix = aref ? wb : wa;
    const float cpen = std::max(0.0f, mp);
    int k = 0;
    for (int i = 0; i < 8; ++i) {
      const V c{(i & 1) ? ih.x : -ih.x, (i & 2) ? ih.y : -ih.y, (i & 4) ? ih.z : -ih.z};
      const V wc = tpoint(ix, c);
      const V lc = tpoint(tinv(rx), wc);
      if (std::fabs(lc.x) <= rh.x + 0.01f && std::fabs(lc.y) <= rh.y + 0.01f && std::fabs(lc.z) <= rh.z + 0.01f) {
        const float off = 0.5f * cpen;
        out.push_back(mk(aref ? wc + best * off : wc - best * off, best, cpen, a.body, b.body, fr, re));
        if (++k >= 4) break;
      }
    }
    if (out.size() == n0) out.push_back(mk(wa.p + d * 0.5f, best, mp, a.body, b.body, fr, re));
  } else {
    out.push_back(mk(wa.p + d * 0.5f, best, mp, a.body, b.body, fr, re));
  }
  return true;
}

// narrowphase.cpp:384-432
bool collide(const Shape& a, const T& ta, const Shape& b, const T& tb, std::vector<Contact>& out) {
  const int A = a.type, B = b.type;
  if (A == 1 && B == 1) return c_sphere_sphere(a, ta, b, tb, out);
  if (A == 1 && B == 2) return c_sphere_plane(a, ta, b, out);
  if (A == 2 && B == 1) return c_sphere_plane(b, tb, a, out);
  if (A == 0 && B == 1) return c_box_sphere(a, ta, b, tb, out);
  if (A == 1 && B == 0) {
    const bool r = c_box_sphere(b, tb, a, ta, out);
    for (auto& c : out) {
      c.n = -c.n;
      std::swap(c.a, c.b);
    }
    return r;
  }
  if (A == 0 && B == 2) return c_box_plane(a, ta, b, out);
  if (A == 2 && B == 0) return c_box_plane(b, tb, a, out);
  if (A == 0 && B == 0) return c_box_box(a, ta, b, tb, out);
  return false;
}

Shape load_shape(const orc_rshape& s) {
  Shape r;
  r.type = s.type;
  r.body = s.body;
  r.half = {s.half[0], s.half[1], s.half[2]};
  r.radius = s.radius;
  r.normal = {s.normal[0], s.normal[1], s.normal[2]};
  r.offset = s.offset;
  r.friction = s.friction;
  r.restitution = s.restitution;
  r.local = make_t({s.local_pos[0], s.local_pos[1], s.local_pos[2]},
                   {s.local_rot[0], s.local_rot[1], s.local_rot[2], s.local_rot[3]});
  return r;
}

T load_pose(const float* p7) { return T{{p7[0], p7[1], p7[2]}, {p7[3], p7[4], p7[5], p7[6]}}; }

void put_contact(const Contact& c, orc_contact& o) {
  o.position[0] = c.pos.x;
  o.position[1] = c.pos.y;
  o.position[2] = c.pos.z;
  o.normal[0] = c.n.x;
  o.normal[1] = c.n.y;
  o.normal[2] = c.n.z;
  o.penetration = c.pen;
  o.body_a = c.a;
  o.body_b = c.b;
  o.friction = c.fr;
  o.restitution = c.re;
  o.acc_n = c.an;
  o.acc_t1 = c.at1;
  o.acc_t2 = c.at2;
}

}  // namespace orr

extern "C" {

int32_t orc_collide_shapes(const orc_rshape* a, const float* pose_a7, const orc_rshape* b, const float* pose_b7,
                           orc_contact* out, int32_t cap) {
  using namespace orr;
  std::vector<Contact> c;
  const bool hit = collide(load_shape(*a), load_pose(pose_a7), load_shape(*b), load_pose(pose_b7), c);
  for (size_t i = 0; i < c.size() && (int32_t)i < cap; ++i) put_contact(c[i], out[i]);
  return hit ? (int32_t)c.size() : 0;
}

int32_t orc_sap_pairs(int32_t n, const float* lo3, const float* hi3, const uint8_t* stat, int32_t* pairs2,
                      int32_t cap) {
  using namespace orr;
  std::vector<Box> bx(n);
  std::vector<char> st(n);
  for (int i = 0; i < n; ++i) {
    bx[i].lo = {lo3[3 * i], lo3[3 * i + 1], lo3[3 * i + 2]};
    bx[i].hi = {hi3[3 * i], hi3[3 * i + 1], hi3[3 * i + 2]};
    st[i] = stat[i] ? 1 : 0;
  }
  const auto pr = sap(bx, st);
  for (size_t k = 0; k < pr.size() && (int32_t)k < cap; ++k) {
    pairs2[2 * k] = pr[k].first;
    pairs2[2 * k + 1] = pr[k].second;
  }
  return (int32_t)pr.size();
}

// One World::step_rigid_pipeline (world.cpp:30-182, sleeping off).  pose7
// (position, quaternion w x y z), lin3, ang3 are updated; force3 / torque3
// are read then cleared.  The step's contacts (after the solve, with their
// accumulated impulses) go to `contacts` (up to cap); their count and the
// broadphase pair count are returned through ncontacts / npairs.
int32_t orc_rigid_step(const orc_rbody* bodies, int32_t nb, const orc_rshape* shapes, int32_t ns,
                       const orc_rsettings* st, const float* gravity3, float dt, float* pose7, float* lin3, float* ang3,
                       float* force3, float* torque3, orc_contact* contacts, int32_t cap, int32_t* ncontacts,
                       int32_t* npairs) {
  using namespace orr;
  auto is_static = [&](int b) { return bodies[b].mass <= 0.0f; };
  auto inv_mass = [&](int b) { return (b >= 0 && bodies[b].mass > 0.0f) ? 1.0f / bodies[b].mass : 0.0f; };
  auto inv_inertia = [&](int b) {
    if (bodies[b].mass <= 0.0f) return zero();
    M I;
    for (int k = 0; k < 9; ++k) I.a[k / 3][k % 3] = bodies[b].inertia[k];
    return inverse(I);
  };
  auto getv = [](const float* a, int i) { return V{a[3 * i], a[3 * i + 1], a[3 * i + 2]}; };
  auto putv = [](float* a, int i, V v) {
    a[3 * i] = v.x;
    a[3 * i + 1] = v.y;
    a[3 * i + 2] = v.z;
  };
  std::vector<T> tf(nb);
  std::vector<V> lv(nb), av(nb);
  for (int i = 0; i < nb; ++i) {
    tf[i] = load_pose(pose7 + 7 * i);
    lv[i] = getv(lin3, i);
    av[i] = getv(ang3, i);
  }
  const V g{gravity3[0], gravity3[1], gravity3[2]};
  // integrate velocity (integrator.cpp:13-24)
  for (int i = 0; i < nb; ++i) {
    if (is_static(i)) continue;
    lv[i] = lv[i] + (g + getv(force3, i) * inv_mass(i)) * dt;
    av[i] = av[i] + mul(inv_inertia(i), getv(torque3, i)) * dt;
    av[i] = av[i] * 0.999f;
  }
  // broadphase
  std::vector<Shape> sh(ns);
  std::vector<Box> bx(ns);
  std::vector<char> stt(ns);
  for (int i = 0; i < ns; ++i) {
    sh[i] = load_shape(shapes[i]);
    if (sh[i].body >= 0) {
      bx[i] = shape_box(sh[i], tf[sh[i].body]);
      stt[i] = is_static(sh[i].body) ? 1 : 0;
    } else {
      bx[i] = shape_box(sh[i], T{});
      stt[i] = 1;
    }
  }
  const auto pr = sap(bx, stt);
  // narrowphase
  std::vector<Contact> C;
  for (const auto& p : pr) {
    const Shape& A = sh[p.first];
    const Shape& B = sh[p.second];
    const T ta = A.body >= 0 ? tf[A.body] : T{};
    const T tb = B.body >= 0 ? tf[B.body] : T{};
    std::vector<Contact> fresh;
    if (collide(A, ta, B, tb, fresh))
      for (auto c : fresh) {
        c.fr = cfr(A.friction, B.friction);
        c.re = cre(A.restitution, B.restitution);
        C.push_back(c);
      }
  }
  // FreeBodySolver::solve (free_body_solver.cpp:30-330)
  if (!C.empty()) {
    struct CD {
      V ra, rb, t1, t2;
      float mn, mt1, mt2, bias;
      M Ia, Ib;
    };
    std::vector<CD> cd(C.size());
    auto vel = [&](int b, V r) { return b >= 0 ? lv[b] + cross(av[b], r) : V{0, 0, 0}; };
    auto dyn = [&](int b) { return b >= 0 && !is_static(b); };
    for (size_t i = 0; i < C.size(); ++i) {
      Contact& c = C[i];
      CD& d = cd[i];
      d.ra = c.a >= 0 ? c.pos - tf[c.a].p : V{0, 0, 0};
      d.rb = c.b >= 0 ? c.pos - tf[c.b].p : V{0, 0, 0};
      const float ima = inv_mass(c.a), imb = inv_mass(c.b);
      auto wI = [&](int b) {
        if (!dyn(b)) return zero();
        const M R = qmatrix(tf[b].q);
        return mul(mul(R, inv_inertia(b)), transpose(R));
      };
      d.Ia = wI(c.a);
      d.Ib = wI(c.b);
      auto em = [&](V dir) {
        const V x = cross(d.ra, dir), y = cross(d.rb, dir);
        float k = ima + imb + dot(x, mul(d.Ia, x)) + dot(y, mul(d.Ib, y));
        if (k > 0) k = 1.0f / k;
        return k;
      };
      d.mn = em(c.n);
      d.t1 = normalized(std::fabs(c.n.x) < 0.9f ? cross(c.n, V{1, 0, 0}) : cross(c.n, V{0, 1, 0}));
      d.t2 = cross(c.n, d.t1);
      d.mt1 = em(d.t1);
      d.mt2 = em(d.t2);
      d.bias = st->baumgarte * std::max(0.0f, c.pen - st->slop) / dt;
      const float rv = dot(c.n, vel(c.b, d.rb) - vel(c.a, d.ra));
      if (rv < -1.0f) d.bias += c.re * rv;
      if (!st->warm_starting) c.an = c.at1 = c.at2 = 0.0f;
    }
    auto push = [&](const Contact& c, const CD& d, V imp) {
      if (dyn(c.a)) {
        lv[c.a] = lv[c.a] - imp * inv_mass(c.a);
        av[c.a] = av[c.a] - mul(d.Ia, cross(d.ra, imp));
      }
      if (dyn(c.b)) {
        lv[c.b] = lv[c.b] + imp * inv_mass(c.b);
        av[c.b] = av[c.b] + mul(d.Ib, cross(d.rb, imp));
      }
    };
    if (st->warm_starting)
      for (size_t i = 0; i < C.size(); ++i) {
        if (cd[i].mn == 0.0f) continue;
        push(C[i], cd[i], C[i].n * C[i].an + cd[i].t1 * C[i].at1 + cd[i].t2 * C[i].at2);
      }
    for (int it = 0; it < st->velocity_iterations; ++it)
      for (size_t i = 0; i < C.size(); ++i) {
        Contact& c = C[i];
        const CD& d = cd[i];
        if (d.mn == 0.0f) continue;
        const float vn = dot(vel(c.b, d.rb) - vel(c.a, d.ra), c.n);
        const float old = c.an;
        c.an = std::max(0.0f, old + d.mn * (-vn + d.bias));
        push(c, d, c.n * (c.an - old));
        const float mf = c.fr * c.an;
        const float v1 = dot(vel(c.b, d.rb) - vel(c.a, d.ra), d.t1);
        const float o1 = c.at1;
        c.at1 = std::max(-mf, std::min(mf, o1 + d.mt1 * (-v1)));
        push(c, d, d.t1 * (c.at1 - o1));
        const float v2 = dot(vel(c.b, d.rb) - vel(c.a, d.ra), d.t2);
        const float o2 = c.at2;
        c.at2 = std::max(-mf, std::min(mf, o2 + d.mt2 * (-v2)));
        push(c, d, d.t2 * (c.at2 - o2));
      }
  }
  // integrate position (integrator.cpp:26-60), clear forces
  for (int i = 0; i < nb; ++i) {
    if (!is_static(i)) {
      const V w = av[i];
      tf[i].p = tf[i].p + lv[i] * dt;
      const Q q = tf[i].q;
      const float dw = 0.5f * (-w.x * q.x - w.y * q.y - w.z * q.z);
      const float dx = 0.5f * (w.x * q.w + w.z * q.y - w.y * q.z);
      const float dy = 0.5f * (w.y * q.w - w.z * q.x + w.x * q.z);
      const float dz = 0.5f * (w.z * q.w + w.y * q.x - w.x * q.y);
      tf[i].q = qnormalized({q.w + dw * dt, q.x + dx * dt, q.y + dy * dt, q.z + dz * dt});
    }
    const float p7[7] = {tf[i].p.x, tf[i].p.y, tf[i].p.z, tf[i].q.w, tf[i].q.x, tf[i].q.y, tf[i].q.z};
    std::copy(p7, p7 + 7, pose7 + 7 * i);
    putv(lin3, i, lv[i]);
    putv(ang3, i, av[i]);
    putv(force3, i, {0, 0, 0});
    putv(torque3, i, {0, 0, 0});
  }
  for (size_t i = 0; i < C.size() && (int32_t)i < cap; ++i) put_contact(C[i], contacts[i]);
  if (ncontacts) *ncontacts = (int32_t)C.size();
  if (npairs) *npairs = (int32_t)pr.size();
  return 0;
}

}  // extern "C"
