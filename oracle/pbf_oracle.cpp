// pbf_oracle.cpp -- CPU restatement of NovaPhy's PBF fluid path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP path
// and the timed CPU baseline in bench.py.  Only tests/, __graft_entry__.smoke()
// and bench.py's CPU-baseline leg may load the library built from it; the
// product (novaphy_amd/) never links, imports or falls back to it.
//
// It restates, in the reference's own operation order and with the reference's
// data structures (std::unordered_map hash grid keyed by the Teschner hash,
// per-particle candidate vectors, sequential float accumulation), the files:
//   include/novaphy/fluid/sph_kernel.h:24-60        poly6 / spiky_grad / spiky
//   src/fluid/neighbor_search.cpp:14-105            SpatialHashGrid
//   src/fluid/particle_state.cpp:11-58              init / generate_fluid_block
//   src/fluid/pbf_solver.cpp:18-250                 PBFSolver::step and passes
//   src/fluid/boundary.cpp:15-192                   Akinci samplers, psi, world pos
//   src/fluid/fluid_world.cpp:101-174               boundary density, coupling
//   include/novaphy/math/math_types.h:80            Transform::transform_point
// Everything is a template on the scalar type: T=float is the faithful FP32
// restatement (built with -ffp-contract=off so no FMA changes rounding),
// T=double is the FP64 recompute used to state the FP32 error.
//
// Eigen is not available here; the small vector type below evaluates
// squaredNorm as x*x + (y*y + z*z) (Eigen 3.4's unrolled pairwise redux for a
// 3-vector), norm as sqrt(squaredNorm), cross as (ay*bz-az*by, az*bx-ax*bz,
// ax*by-ay*bx), and quaternion*vector as Eigen's _transformVector
// (uv = 2 q.vec x v; v + w*uv + q.vec x uv).
//
// Vorticity confinement is NOT part of the reference (PBFSettings declares
// vorticity_epsilon, pbf_solver.h:20, and never reads it).  The optional
// vorticity pass below is this project's own definition, shared with the HIP
// path.  There is no reference to compare it with; it is pinned by the
// mathematics instead: uniform flow gives omega = 0, rigid rotation gives
// omega = 2 Omega, and a Gaussian vortex gives the analytic curl, eta =
// grad|omega| toward the axis and a confinement force eps (N x omega) along
// the rotation with |f| = eps |omega| (tests/test_oracle_pins.py).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>

extern "C" {

// Same field order and meaning as PBFSettings (pbf_solver.h:13-38), with the
// C++-only fields (vorticity_epsilon, corr_*) included.
typedef struct orc_params {
  float rest_density;
  float kernel_radius;
  float particle_radius;
  int32_t solver_iterations;
  float epsilon;
  float xsph_viscosity;
  float vorticity_epsilon;
  float corr_delta_q;
  float corr_k;
  int32_t corr_n;
  int32_t use_domain_bounds;
  float domain_lower[3];
  float domain_upper[3];
} orc_params;

// Shape description for boundary sampling (shape.h:43-142).
// type: 0 box (a = half extents), 1 sphere (a[0] = radius),
//       2 plane (a[0..2] = unit normal, a[3] = offset).
// local_rot is a unit quaternion (w, x, y, z).
typedef struct orc_shape {
  int32_t type;
  int32_t body;
  float a[4];
  float local_pos[3];
  float local_rot[4];
} orc_shape;

}  // extern "C"

namespace orc {

template <class T>
struct V3 {
  T x, y, z;
};
template <class T>
inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T>
inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T>
inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T>
inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T>
inline T sqn(V3<T> a) { return a.x * a.x + (a.y * a.y + a.z * a.z); }
template <class T>
inline T norm(V3<T> a) { return std::sqrt(sqn(a)); }
template <class T>
inline V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T>
inline T comp(const V3<T>& a, int d) { return d == 0 ? a.x : (d == 1 ? a.y : a.z); }
template <class T>
inline T& comp(V3<T>& a, int d) { return d == 0 ? a.x : (d == 1 ? a.y : a.z); }

template <class T>
struct Quat {
  T w, x, y, z;
};
// Eigen's Quaternion::_transformVector.
template <class T>
inline V3<T> qrot(Quat<T> q, V3<T> v) {
  V3<T> qv{q.x, q.y, q.z};
  V3<T> uv = cross(qv, v);
  uv = uv + uv;
  V3<T> wuv = q.w * uv;
  V3<T> c = cross(qv, uv);
  return (v + wuv) + c;
}

template <class T>
struct C;
template <>
struct C<float> {
  static constexpr float PI = 3.14159265358979323846f;  // math_utils.h:53
};
template <>
struct C<double> {
  static constexpr double PI = 3.14159265358979323846;
};

// sph_kernel.h:24-30
template <class T>
inline T poly6(T r_sq, T h) {
  if (r_sq >= h * h) return T(0);
  T h2 = h * h;
  T diff = h2 - r_sq;
  T h9 = h2 * h2 * h2 * h2 * h;
  return (T(315) / (T(64) * C<T>::PI * h9)) * diff * diff * diff;
}

// sph_kernel.h:39-46
template <class T>
inline V3<T> spiky_grad(V3<T> r, T h) {
  T r_len = norm(r);
  if (r_len <= T(1e-7f) || r_len >= h) return {T(0), T(0), T(0)};
  T diff = h - r_len;
  T h6 = h * h * h * h * h * h;
  T coeff = -(T(45) / (C<T>::PI * h6)) * diff * diff / r_len;
  return coeff * r;
}

// sph_kernel.h:55-60
template <class T>
inline T spiky(T r_len, T h) {
  if (r_len >= h || r_len < T(0)) return T(0);
  T diff = h - r_len;
  T h6 = h * h * h * h * h * h;
  return (T(15) / (C<T>::PI * h6)) * diff * diff * diff;
}

// sph_kernel.h:69-80
template <class T>
inline T cubic_spline(T r_len, T h) {
  T q = r_len / h;
  if (q >= T(1)) return T(0);
  T h3 = h * h * h;
  T sigma = T(8) / (C<T>::PI * h3);
  if (q <= T(0.5)) return sigma * (T(6) * (q * q * q - q * q) + T(1));
  T t = T(1) - q;
  return sigma * T(2) * t * t * t;
}

// neighbor_search.cpp:89-97
inline uint64_t hash_cell(int cx, int cy, int cz) {
  const uint64_t p1 = 73856093ULL, p2 = 19349663ULL, p3 = 83492791ULL;
  return (static_cast<uint64_t>(cx) * p1) ^ (static_cast<uint64_t>(cy) * p2) ^
         (static_cast<uint64_t>(cz) * p3);
}

// neighbor_search.cpp:14-105
template <class T>
struct HashGrid {
  T cell_size;
  std::unordered_map<uint64_t, std::vector<int>> cells;
  explicit HashGrid(T cs) : cell_size(cs) {}
  void world_to_cell(V3<T> p, int& cx, int& cy, int& cz) const {
    T inv = T(1) / cell_size;
    cx = static_cast<int>(std::floor(p.x * inv));
    cy = static_cast<int>(std::floor(p.y * inv));
    cz = static_cast<int>(std::floor(p.z * inv));
  }
  void build(const std::vector<V3<T>>& pos) {
    cells.clear();
    for (size_t i = 0; i < pos.size(); ++i) {
      int cx, cy, cz;
      world_to_cell(pos[i], cx, cy, cz);
      cells[hash_cell(cx, cy, cz)].push_back(static_cast<int>(i));
    }
  }
  void query(V3<T> p, std::vector<int>& out) const {
    out.clear();
    int cx, cy, cz;
    world_to_cell(p, cx, cy, cz);
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          auto it = cells.find(hash_cell(cx + dx, cy + dy, cz + dz));
          if (it == cells.end()) continue;
          for (int idx : it->second) out.push_back(idx);
        }
  }
};

template <class T>
struct Params {
  T rest_density, kernel_radius, particle_radius;
  int solver_iterations;
  T epsilon, xsph_viscosity, vorticity_epsilon, corr_delta_q, corr_k;
  int corr_n;
  bool use_domain_bounds;
  V3<T> lo, hi;
  static Params from(const orc_params& p) {
    Params q;
    q.rest_density = T(p.rest_density);
    q.kernel_radius = T(p.kernel_radius);
    q.particle_radius = T(p.particle_radius);
    q.solver_iterations = p.solver_iterations;
    q.epsilon = T(p.epsilon);
    q.xsph_viscosity = T(p.xsph_viscosity);
    q.vorticity_epsilon = T(p.vorticity_epsilon);
    q.corr_delta_q = T(p.corr_delta_q);
    q.corr_k = T(p.corr_k);
    q.corr_n = p.corr_n;
    q.use_domain_bounds = p.use_domain_bounds != 0;
    q.lo = {T(p.domain_lower[0]), T(p.domain_lower[1]), T(p.domain_lower[2])};
    q.hi = {T(p.domain_upper[0]), T(p.domain_upper[1]), T(p.domain_upper[2])};
    return q;
  }
};

// ParticleState (particle_state.h:16-44), kept AoS like the reference.
template <class T>
struct State {
  std::vector<V3<T>> x, v, xs, dp;
  std::vector<T> rho, lam;
};

// PBFSolver::step (pbf_solver.cpp:18-127) plus its passes (:129-250).
template <class T>
struct Solver {
  Params<T> s;
  HashGrid<T> grid;
  std::vector<std::vector<int>> nbr;
  std::vector<V3<T>> omega;  // vorticity extension only
  explicit Solver(const Params<T>& p) : s(p), grid(p.kernel_radius) {}

  void density(State<T>& st, T m) {
    const int n = (int)st.x.size();
    const T h = s.kernel_radius;
    for (int i = 0; i < n; ++i) {
      T d = T(0);
      for (int j : nbr[i]) {
        V3<T> r = st.xs[i] - st.xs[j];
        d += m * poly6(sqn(r), h);
      }
      st.rho[i] = d;
    }
  }
  void lambda(State<T>& st, T m) {
    const int n = (int)st.x.size();
    const T h = s.kernel_radius, eps = s.epsilon;
    const T inv_rho0 = T(1) / s.rest_density;
    for (int i = 0; i < n; ++i) {
      T Ci = st.rho[i] * inv_rho0 - T(1);
      if (Ci < T(0)) {
        st.lam[i] = T(0);
        continue;
      }
      T sum_grad_sq = T(0);
      V3<T> gi{T(0), T(0), T(0)};
      for (int j : nbr[i]) {
        if (j == i) continue;
        V3<T> r = st.xs[i] - st.xs[j];
        V3<T> g = spiky_grad(r, h);
        V3<T> gj = (m * inv_rho0) * g;
        sum_grad_sq += sqn(gj);
        gi = gi + gj;
      }
      sum_grad_sq += sqn(gi);
      st.lam[i] = -Ci / (sum_grad_sq + eps);
    }
  }
  void delta(State<T>& st, T m) {
    const int n = (int)st.x.size();
    const T h = s.kernel_radius;
    const T inv_rho0 = T(1) / s.rest_density;
    const T dq = s.corr_delta_q * h;
    const T wdq = poly6(dq * dq, h);
    for (int i = 0; i < n; ++i) {
      V3<T> d{T(0), T(0), T(0)};
      for (int j : nbr[i]) {
        if (j == i) continue;
        V3<T> r = st.xs[i] - st.xs[j];
        T r_sq = sqn(r);
        T ls = st.lam[i] + st.lam[j];
        T sc = T(0);
        if (ls < T(-1e-10f) && wdq > T(1e-10f)) {
          T ratio = poly6(r_sq, h) / wdq;
          ratio = std::min(ratio, T(1));
          T rp = ratio * ratio;
          if (s.corr_n == 4) rp = rp * rp;
          sc = -s.corr_k * rp;
        }
        T coef = (ls + sc) * m;
        d = d + coef * spiky_grad(r, h);
      }
      st.dp[i] = d * inv_rho0;
    }
  }
  void xsph(State<T>& st, T m) {
    const int n = (int)st.x.size();
    const T h = s.kernel_radius, c = s.xsph_viscosity, rho0 = s.rest_density;
    for (int i = 0; i < n; ++i) {
      V3<T> vc{T(0), T(0), T(0)};
      for (int j : nbr[i]) {
        if (j == i) continue;
        V3<T> vij = st.v[j] - st.v[i];
        V3<T> r = st.xs[i] - st.xs[j];
        T w = poly6(sqn(r), h);
        T rj = (st.rho[j] > T(1e-6f)) ? st.rho[j] : rho0;
        vc = vc + ((m / rj) * vij) * w;
      }
      st.dp[i] = vc;
    }
    for (int i = 0; i < n; ++i) st.v[i] = st.v[i] + c * st.dp[i];
  }
  // Vorticity confinement: this project's definition (not in the reference).
  //   omega_i = sum_{j!=i} V_j (v_j - v_i) x gradW(x_j - x_i),  V_j = m / rho~_j
  //   eta_i   = sum_{j!=i} V_j (|omega_j| - |omega_i|) gradW(x_i - x_j)
  //   N_i = eta_i / |eta_i| (0 if |eta_i| <= 1e-9);  f_i = eps_v (N_i x omega_i)
  //   v_i += dt * f_i     (applied after XSPH, before the domain clamp)
  void vorticity_omega(State<T>& st, T m) {
    const int n = (int)st.x.size();
    const T h = s.kernel_radius, rho0 = s.rest_density;
    omega.assign(n, {T(0), T(0), T(0)});
    for (int i = 0; i < n; ++i) {
      V3<T> w{T(0), T(0), T(0)};
      for (int j : nbr[i]) {
        if (j == i) continue;
        V3<T> vij = st.v[j] - st.v[i];
        V3<T> g = spiky_grad(st.xs[j] - st.xs[i], h);
        T rj = (st.rho[j] > T(1e-6f)) ? st.rho[j] : rho0;
        w = w + (m / rj) * cross(vij, g);
      }
      omega[i] = w;
    }
  }
  // Confinement force of particle i from the omega field: eta_i = grad|omega|
  // (SPH gradient), N = eta / |eta|, f = eps_v (N x omega_i).  eta_out gets eta_i.
  V3<T> vorticity_force(const State<T>& st, T m, int i, V3<T>* eta_out) const {
    const T h = s.kernel_radius, rho0 = s.rest_density, ev = s.vorticity_epsilon;
    T wi = norm(omega[i]);
    V3<T> eta{T(0), T(0), T(0)};
    for (int j : nbr[i]) {
      if (j == i) continue;
      V3<T> g = spiky_grad(st.xs[i] - st.xs[j], h);
      T rj = (st.rho[j] > T(1e-6f)) ? st.rho[j] : rho0;
      eta = eta + ((m / rj) * (norm(omega[j]) - wi)) * g;
    }
    if (eta_out) *eta_out = eta;
    T el = norm(eta);
    V3<T> f{T(0), T(0), T(0)};
    if (el > T(1e-9f)) {
      V3<T> N = eta * (T(1) / el);
      f = ev * cross(N, omega[i]);
    }
    return f;
  }
  void vorticity_apply(State<T>& st, T m, T dt, std::vector<V3<T>>& vnew) {
    const int n = (int)st.x.size();
    for (int i = 0; i < n; ++i) vnew[i] = st.v[i] + dt * vorticity_force(st, m, i, nullptr);
  }

  void step(State<T>& st, T dt, V3<T> g, T m, bool skip_grid = false) {
    const int n = (int)st.x.size();
    if (n == 0) return;
    const T h = s.kernel_radius;
    grid.cell_size = h;
    for (int i = 0; i < n; ++i) {
      st.v[i] = st.v[i] + g * dt;
      st.xs[i] = st.x[i] + st.v[i] * dt;
    }
    if (!skip_grid) {
      grid.build(st.xs);
      nbr.resize(n);
      std::vector<int> scratch;
      for (int i = 0; i < n; ++i) {
        grid.query(st.xs[i], scratch);
        nbr[i] = scratch;
      }
    }
    for (int it = 0; it < s.solver_iterations; ++it) {
      density(st, m);
      lambda(st, m);
      delta(st, m);
      for (int i = 0; i < n; ++i) st.xs[i] = st.xs[i] + st.dp[i];
    }
    const T inv_dt = T(1) / dt;
    const T max_speed = h / dt;
    const T max_speed_sq = max_speed * max_speed;
    for (int i = 0; i < n; ++i) {
      st.v[i] = (st.xs[i] - st.x[i]) * inv_dt;
      T sp = sqn(st.v[i]);
      if (sp > max_speed_sq) st.v[i] = st.v[i] * (max_speed / std::sqrt(sp));
    }
    const bool vort = s.vorticity_epsilon != T(0);
    if (vort) vorticity_omega(st, m);
    if (s.xsph_viscosity > T(0)) xsph(st, m);
    if (vort) {
      std::vector<V3<T>> vnew(n);
      vorticity_apply(st, m, dt, vnew);
      st.v.swap(vnew);
    }
    if (s.use_domain_bounds) {
      const T pr = s.particle_radius;
      V3<T> lo = s.lo + V3<T>{pr, pr, pr};
      V3<T> hi = s.hi - V3<T>{pr, pr, pr};
      for (int i = 0; i < n; ++i) {
        V3<T>& p = st.xs[i];
        for (int d = 0; d < 3; ++d) {
          if (comp(p, d) < comp(lo, d)) {
            comp(p, d) = comp(lo, d);
            comp(st.v[i], d) = T(0);
          } else if (comp(p, d) > comp(hi, d)) {
            comp(p, d) = comp(hi, d);
            comp(st.v[i], d) = T(0);
          }
        }
        st.x[i] = p;
      }
    } else {
      for (int i = 0; i < n; ++i) st.x[i] = st.xs[i];
    }
  }
};

template <class T>
void load_state(State<T>& st, int n, const T* pos, const T* vel, const T* rho, const T* lam) {
  st.x.resize(n);
  st.v.resize(n);
  st.xs.resize(n);
  st.dp.assign(n, {T(0), T(0), T(0)});
  st.rho.assign(n, T(0));
  st.lam.assign(n, T(0));
  for (int i = 0; i < n; ++i) {
    st.x[i] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    st.v[i] = vel ? V3<T>{vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]} : V3<T>{T(0), T(0), T(0)};
    if (rho) st.rho[i] = rho[i];
    if (lam) st.lam[i] = lam[i];
  }
  st.xs = st.x;  // ParticleState::init: predicted_positions = positions
}
template <class T>
void store_state(const State<T>& st, T* pos, T* vel, T* rho, T* lam) {
  const int n = (int)st.x.size();
  for (int i = 0; i < n; ++i) {
    if (pos) {
      pos[3 * i] = st.x[i].x;
      pos[3 * i + 1] = st.x[i].y;
      pos[3 * i + 2] = st.x[i].z;
    }
    if (vel) {
      vel[3 * i] = st.v[i].x;
      vel[3 * i + 1] = st.v[i].y;
      vel[3 * i + 2] = st.v[i].z;
    }
    if (rho) rho[i] = st.rho[i];
    if (lam) lam[i] = st.lam[i];
  }
}

// ---------------------------------------------------------------- boundary
// boundary.cpp:178-192 (+ Transform::transform_point, math_types.h:80)
template <class T>
std::vector<V3<T>> boundary_world(int nb, const T* local3, const int32_t* body, int nbodies,
                                  const T* pose7) {
  std::vector<V3<T>> out(nb);
  for (int b = 0; b < nb; ++b) {
    V3<T> p{local3[3 * b], local3[3 * b + 1], local3[3 * b + 2]};
    int bi = body[b];
    if (bi >= 0 && bi < nbodies) {
      const T* P = pose7 + 7 * bi;
      V3<T> pos{P[0], P[1], P[2]};
      Quat<T> q{P[3], P[4], P[5], P[6]};
      out[b] = pos + qrot(q, p);
    } else {
      out[b] = p;
    }
  }
  return out;
}

// boundary.cpp:152-176
template <class T>
void boundary_volumes(const std::vector<V3<T>>& wp, T h, T* psi) {
  const int n = (int)wp.size();
  if (n == 0) return;
  HashGrid<T> grid(h);
  grid.build(wp);
  std::vector<int> scratch;
  for (int i = 0; i < n; ++i) {
    T sum = T(0);
    grid.query(wp[i], scratch);
    for (int j : scratch) sum += poly6(sqn(wp[i] - wp[j]), h);
    psi[i] = (sum > T(1e-10f)) ? T(1) / sum : T(0);
  }
}

// fluid_world.cpp:101-123
template <class T>
void boundary_density(int n, const T* pos3, T* rho, const std::vector<V3<T>>& bp, const T* psi,
                      T h, T rho0) {
  HashGrid<T> grid(h);
  grid.build(bp);
  std::vector<int> scratch;
  for (int i = 0; i < n; ++i) {
    V3<T> xi{pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2]};
    T bd = T(0);
    grid.query(xi, scratch);
    for (int j : scratch) {
      T r_sq = sqn(xi - bp[j]);
      bd += rho0 * psi[j] * poly6(r_sq, h);
    }
    rho[i] += bd;
  }
}

// fluid_world.cpp:125-174.  Per-body force / torque accumulated in boundary
// particle order, as World::apply_force / apply_torque do (state.cpp:63-78).
template <class T>
void coupling(int n, const T* pos3, const T* rho, const std::vector<V3<T>>& bp, const T* psi,
              const int32_t* body, int nbodies, const T* pose7, const uint8_t* is_static, T h,
              T rho0, T m, T* force3, T* torque3, T* per_particle3) {
  std::vector<V3<T>> fx(n);
  for (int i = 0; i < n; ++i) fx[i] = {pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2]};
  HashGrid<T> grid(h);
  grid.build(fx);
  for (int k = 0; k < 3 * nbodies; ++k) {
    force3[k] = T(0);
    torque3[k] = T(0);
  }
  std::vector<int> scratch;
  const int nb = (int)bp.size();
  for (int b = 0; b < nb; ++b) {
    if (per_particle3) per_particle3[3 * b] = per_particle3[3 * b + 1] = per_particle3[3 * b + 2] = T(0);
    int bi = body[b];
    if (bi < 0) continue;
    if (is_static[bi]) continue;
    V3<T> f{T(0), T(0), T(0)};
    T psi_b = psi[b];
    grid.query(bp[b], scratch);
    for (int j : scratch) {
      V3<T> r = bp[b] - fx[j];
      T r_sq = sqn(r);
      if (r_sq >= h * h) continue;
      T rho_f = std::max(rho[j], T(1));
      T pressure = std::max(T(0), rho_f - rho0);
      T pc = m * pressure / (rho_f * rho_f);
      V3<T> g = spiky_grad(r, h);
      f = f + ((pc * rho0) * psi_b) * g;
    }
    if (per_particle3) {
      per_particle3[3 * b] = f.x;
      per_particle3[3 * b + 1] = f.y;
      per_particle3[3 * b + 2] = f.z;
    }
    if (sqn(f) > T(1e-20f)) {
      force3[3 * bi] += f.x;
      force3[3 * bi + 1] += f.y;
      force3[3 * bi + 2] += f.z;
      V3<T> com{pose7[7 * bi], pose7[7 * bi + 1], pose7[7 * bi + 2]};
      V3<T> tq = cross(bp[b] - com, f);
      torque3[3 * bi] += tq.x;
      torque3[3 * bi + 1] += tq.y;
      torque3[3 * bi + 2] += tq.z;
    }
  }
}

}  // namespace orc

using namespace orc;

// ---------------------------------------------------------------------------
// C ABI for ctypes (tests / bench CPU baseline only).
extern "C" {

float orc_poly6(float r_sq, float h) { return poly6<float>(r_sq, h); }
double orc_poly6_f64(double r_sq, double h) { return poly6<double>(r_sq, h); }
void orc_spiky_grad(const float* r, float h, float* out) {
  V3<float> g = spiky_grad<float>({r[0], r[1], r[2]}, h);
  out[0] = g.x;
  out[1] = g.y;
  out[2] = g.z;
}
float orc_spiky(float r_len, float h) { return spiky<float>(r_len, h); }
float orc_cubic_spline(float r_len, float h) { return cubic_spline<float>(r_len, h); }
uint64_t orc_hash_cell(int cx, int cy, int cz) { return hash_cell(cx, cy, cz); }

// particle_state.cpp:31-58 -- returns the count; writes positions if out != NULL.
int32_t orc_generate_fluid_block(const float* lower, const float* upper, float spacing, float* out) {
  if (spacing <= 0.0f) return 0;
  int nx = std::max(1, static_cast<int>(std::floor((upper[0] - lower[0]) / spacing)) + 1);
  int ny = std::max(1, static_cast<int>(std::floor((upper[1] - lower[1]) / spacing)) + 1);
  int nz = std::max(1, static_cast<int>(std::floor((upper[2] - lower[2]) / spacing)) + 1);
  if (out) {
    int k = 0;
    for (int iz = 0; iz < nz; ++iz)
      for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
          out[3 * k] = lower[0] + ix * spacing;
          out[3 * k + 1] = lower[1] + iy * spacing;
          out[3 * k + 2] = lower[2] + iz * spacing;
          ++k;
        }
  }
  return nx * ny * nz;
}

// Grid diagnostics on one build (neighbor_search.cpp:14-49):
//  keys[i]      Teschner key of particle i's cell
//  cand_cnt[i]  length of query_neighbors(x_i) (duplicates included)
//  nbr_cnt[i]   candidates j != i with squaredNorm(x_i - x_j) < radius^2
//               (duplicates included, as the PBF loops count them)
//  cand_list    optional flattened candidate lists (size = returned total)
int64_t orc_grid_debug(int32_t n, const float* pos3, float cell, float radius, uint64_t* keys,
                       int32_t* cand_cnt, int32_t* nbr_cnt, int32_t* cand_list,
                       int64_t cand_list_cap) {
  std::vector<V3<float>> p(n);
  for (int i = 0; i < n; ++i) p[i] = {pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2]};
  HashGrid<float> g(cell);
  g.build(p);
  std::vector<int> scratch;
  int64_t total = 0;
  const float r2 = radius * radius;
  for (int i = 0; i < n; ++i) {
    int cx, cy, cz;
    g.world_to_cell(p[i], cx, cy, cz);
    if (keys) keys[i] = hash_cell(cx, cy, cz);
    g.query(p[i], scratch);
    if (cand_cnt) cand_cnt[i] = (int32_t)scratch.size();
    int k = 0;
    for (int j : scratch) {
      if (cand_list && total < cand_list_cap) cand_list[total] = j;
      ++total;
      if (j != i && sqn(p[i] - p[j]) < r2) ++k;
    }
    if (nbr_cnt) nbr_cnt[i] = k;
  }
  return total;
}

// query_neighbors for arbitrary query points against a built point set.
int64_t orc_grid_query(int32_t n, const float* pos3, float cell, int32_t nq, const float* q3,
                       int32_t* counts, int32_t* list, int64_t cap) {
  std::vector<V3<float>> p(n);
  for (int i = 0; i < n; ++i) p[i] = {pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2]};
  HashGrid<float> g(cell);
  g.build(p);
  std::vector<int> scratch;
  int64_t total = 0;
  for (int i = 0; i < nq; ++i) {
    g.query({q3[3 * i], q3[3 * i + 1], q3[3 * i + 2]}, scratch);
    if (counts) counts[i] = (int32_t)scratch.size();
    for (int j : scratch) {
      if (list && total < cap) list[total] = j;
      ++total;
    }
  }
  return total;
}

// neighbor_search.cpp:57-87 -- pairs (i<j) with squaredNorm < radius^2.
int64_t orc_grid_all_pairs(int32_t n, const float* pos3, float cell, float radius, int32_t* pairs,
                           int64_t cap) {
  std::vector<V3<float>> p(n);
  for (int i = 0; i < n; ++i) p[i] = {pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2]};
  HashGrid<float> g(cell);
  g.build(p);
  const float r2 = radius * radius;
  int64_t total = 0;
  for (int i = 0; i < n; ++i) {
    int cx, cy, cz;
    g.world_to_cell(p[i], cx, cy, cz);
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          auto it = g.cells.find(hash_cell(cx + dx, cy + dy, cz + dz));
          if (it == g.cells.end()) continue;
          for (int j : it->second) {
            if (j <= i) continue;
            if (sqn(p[i] - p[j]) < r2) {
              if (pairs && total < cap) {
                pairs[2 * total] = i;
                pairs[2 * total + 1] = j;
              }
              ++total;
            }
          }
        }
  }
  return total;
}

// Run nsteps of PBFSolver::step on an AoS state (in/out).  FP32 faithful.
int32_t orc_pbf_run_f32(const orc_params* prm, int32_t n, float* pos3, float* vel3, float* rho,
                        float* lam, float dt, const float* g3, float mass, int32_t nsteps) {
  State<float> st;
  load_state<float>(st, n, pos3, vel3, nullptr, nullptr);
  Solver<float> sv(Params<float>::from(*prm));
  for (int s = 0; s < nsteps; ++s) sv.step(st, dt, {g3[0], g3[1], g3[2]}, mass);
  store_state<float>(st, pos3, vel3, rho, lam);
  return 0;
}

// FP64 recompute of the same steps (inputs widened from float).
int32_t orc_pbf_run_f64(const orc_params* prm, int32_t n, double* pos3, double* vel3, double* rho,
                        double* lam, double dt, const double* g3, double mass, int32_t nsteps) {
  State<double> st;
  load_state<double>(st, n, pos3, vel3, nullptr, nullptr);
  Solver<double> sv(Params<double>::from(*prm));
  for (int s = 0; s < nsteps; ++s) sv.step(st, dt, {g3[0], g3[1], g3[2]}, mass);
  store_state<double>(st, pos3, vel3, rho, lam);
  return 0;
}

// Persistent solver handle for timing (bench CPU baseline): avoids
// re-allocating the state between timed steps.
struct orc_sim {
  State<float> st;
  Solver<float> sv;
  orc_sim(const Params<float>& p) : sv(p) {}
};
void* orc_sim_create(const orc_params* prm, int32_t n, const float* pos3, const float* vel3) {
  orc_sim* s = new orc_sim(Params<float>::from(*prm));
  load_state<float>(s->st, n, pos3, vel3, nullptr, nullptr);
  return s;
}
int32_t orc_sim_step(void* h, float dt, const float* g3, float mass, int32_t nsteps) {
  orc_sim* s = (orc_sim*)h;
  for (int k = 0; k < nsteps; ++k) s->sv.step(s->st, dt, {g3[0], g3[1], g3[2]}, mass);
  return 0;
}
void orc_sim_get(void* h, float* pos3, float* vel3, float* rho, float* lam) {
  store_state<float>(((orc_sim*)h)->st, pos3, vel3, rho, lam);
}
void orc_sim_destroy(void* h) { delete (orc_sim*)h; }

// ---- boundary sampling (boundary.cpp:15-150), FP32 only (setup path) ----
static V3<float> xform_point(const orc_shape& s, V3<float> p) {
  // Transform(p, q) normalises q on construction (math_types.h:50).
  float w = s.local_rot[0], x = s.local_rot[1], y = s.local_rot[2], z = s.local_rot[3];
  float nrm = std::sqrt(w * w + (x * x + (y * y + z * z)));
  Quat<float> q{w / nrm, x / nrm, y / nrm, z / nrm};
  V3<float> t{s.local_pos[0], s.local_pos[1], s.local_pos[2]};
  return t + qrot(q, p);
}
static V3<float> normalized(V3<float> v) {
  float z = sqn(v);
  if (z > 0.0f) {
    float l = std::sqrt(z);
    return {v.x / l, v.y / l, v.z / l};
  }
  return v;
}

// Returns the number of boundary particles; fills local3/body if non-NULL.
int32_t orc_sample_boundaries(int32_t nshapes, const orc_shape* shapes, float spacing,
                              float plane_extent, float* local3, int32_t* body) {
  int32_t k = 0;
  auto emit = [&](V3<float> p, int b) {
    if (local3) {
      local3[3 * k] = p.x;
      local3[3 * k + 1] = p.y;
      local3[3 * k + 2] = p.z;
    }
    if (body) body[k] = b;
    ++k;
  };
  for (int si = 0; si < nshapes; ++si) {
    const orc_shape& s = shapes[si];
    if (s.type == 0) {
      const float he[3] = {s.a[0], s.a[1], s.a[2]};
      for (int axis = 0; axis < 3; ++axis) {
        int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
        int n1 = std::max(1, static_cast<int>(std::floor(2.0f * he[a1] / spacing)) + 1);
        int n2 = std::max(1, static_cast<int>(std::floor(2.0f * he[a2] / spacing)) + 1);
        for (int sign = -1; sign <= 1; sign += 2)
          for (int i1 = 0; i1 < n1; ++i1)
            for (int i2 = 0; i2 < n2; ++i2) {
              V3<float> lp{0.0f, 0.0f, 0.0f};
              comp(lp, axis) = sign * he[axis];
              comp(lp, a1) = -he[a1] + i1 * spacing;
              comp(lp, a2) = -he[a2] + i2 * spacing;
              emit(xform_point(s, lp), s.body);
            }
      }
    } else if (s.type == 1) {
      const float r = s.a[0];
      const float pi = 3.14159265358979323846f;  // std::numbers::pi_v<float>
      const float area = 4.0f * pi * r * r;
      int n = std::max(12, static_cast<int>(area / (spacing * spacing)));
      const float golden_ratio = (1.0f + std::sqrt(5.0f)) / 2.0f;
      const float golden_angle = 2.0f * pi / (golden_ratio * golden_ratio);
      for (int i = 0; i < n; ++i) {
        float theta = std::acos(1.0f - 2.0f * (i + 0.5f) / n);
        float phi = golden_angle * i;
        V3<float> lp{r * std::sin(theta) * std::cos(phi), r * std::sin(theta) * std::sin(phi),
                     r * std::cos(theta)};
        emit(xform_point(s, lp), s.body);
      }
    } else if (s.type == 2) {
      V3<float> nrm{s.a[0], s.a[1], s.a[2]};
      float off = s.a[3];
      V3<float> t1, t2;
      if (std::abs(nrm.y) < 0.9f)
        t1 = normalized(cross(V3<float>{0, 1, 0}, nrm));
      else
        t1 = normalized(cross(V3<float>{1, 0, 0}, nrm));
      t2 = normalized(cross(nrm, t1));
      int n = static_cast<int>(std::floor(2.0f * plane_extent / spacing)) + 1;
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
          float ss = -plane_extent + i * spacing;
          float tt = -plane_extent + j * spacing;
          V3<float> p = (nrm * off + t1 * ss) + t2 * tt;
          emit(p, s.body);
        }
    }
  }
  return k;
}

int32_t orc_boundary_world_positions(int32_t nb, const float* local3, const int32_t* body,
                                     int32_t nbodies, const float* pose7, float* out3) {
  auto w = boundary_world<float>(nb, local3, body, nbodies, pose7);
  for (int b = 0; b < nb; ++b) {
    out3[3 * b] = w[b].x;
    out3[3 * b + 1] = w[b].y;
    out3[3 * b + 2] = w[b].z;
  }
  return 0;
}

int32_t orc_boundary_volumes(int32_t nb, const float* wpos3, float h, float* psi) {
  std::vector<V3<float>> w(nb);
  for (int b = 0; b < nb; ++b) w[b] = {wpos3[3 * b], wpos3[3 * b + 1], wpos3[3 * b + 2]};
  boundary_volumes<float>(w, h, psi);
  return 0;
}
int32_t orc_boundary_volumes_f64(int32_t nb, const double* wpos3, double h, double* psi) {
  std::vector<V3<double>> w(nb);
  for (int b = 0; b < nb; ++b) w[b] = {wpos3[3 * b], wpos3[3 * b + 1], wpos3[3 * b + 2]};
  boundary_volumes<double>(w, h, psi);
  return 0;
}

int32_t orc_boundary_density(int32_t n, const float* pos3, float* rho, int32_t nb,
                             const float* bpos3, const float* psi, float h, float rho0) {
  std::vector<V3<float>> bp(nb);
  for (int b = 0; b < nb; ++b) bp[b] = {bpos3[3 * b], bpos3[3 * b + 1], bpos3[3 * b + 2]};
  boundary_density<float>(n, pos3, rho, bp, psi, h, rho0);
  return 0;
}
int32_t orc_boundary_density_f64(int32_t n, const double* pos3, double* rho, int32_t nb,
                                 const double* bpos3, const double* psi, double h, double rho0) {
  std::vector<V3<double>> bp(nb);
  for (int b = 0; b < nb; ++b) bp[b] = {bpos3[3 * b], bpos3[3 * b + 1], bpos3[3 * b + 2]};
  boundary_density<double>(n, pos3, rho, bp, psi, h, rho0);
  return 0;
}

int32_t orc_coupling(int32_t n, const float* pos3, const float* rho, int32_t nb, const float* bpos3,
                     const float* psi, const int32_t* body, int32_t nbodies, const float* pose7,
                     const uint8_t* is_static, float h, float rho0, float mass, float* force3,
                     float* torque3, float* per_particle3) {
  std::vector<V3<float>> bp(nb);
  for (int b = 0; b < nb; ++b) bp[b] = {bpos3[3 * b], bpos3[3 * b + 1], bpos3[3 * b + 2]};
  coupling<float>(n, pos3, rho, bp, psi, body, nbodies, pose7, is_static, h, rho0, mass, force3,
                  torque3, per_particle3);
  return 0;
}
int32_t orc_coupling_f64(int32_t n, const double* pos3, const double* rho, int32_t nb,
                         const double* bpos3, const double* psi, const int32_t* body,
                         int32_t nbodies, const double* pose7, const uint8_t* is_static, double h,
                         double rho0, double mass, double* force3, double* torque3) {
  std::vector<V3<double>> bp(nb);
  for (int b = 0; b < nb; ++b) bp[b] = {bpos3[3 * b], bpos3[3 * b + 1], bpos3[3 * b + 2]};
  coupling<double>(n, pos3, rho, bp, psi, body, nbodies, pose7, is_static, h, rho0, mass, force3,
                   torque3, nullptr);
  return 0;
}

// Vorticity field only (for the closed-form checks): omega for given state,
// using the full candidate lists of one grid build at x (no prediction).
int32_t orc_vorticity_omega(const orc_params* prm, int32_t n, const float* pos3, const float* vel3,
                            const float* rho, float mass, float* omega3) {
  State<float> st;
  load_state<float>(st, n, pos3, vel3, rho, nullptr);
  Solver<float> sv(Params<float>::from(*prm));
  sv.grid.build(st.xs);
  sv.nbr.resize(n);
  std::vector<int> scratch;
  for (int i = 0; i < n; ++i) {
    sv.grid.query(st.xs[i], scratch);
    sv.nbr[i] = scratch;
  }
  sv.vorticity_omega(st, mass);
  for (int i = 0; i < n; ++i) {
    omega3[3 * i] = sv.omega[i].x;
    omega3[3 * i + 1] = sv.omega[i].y;
    omega3[3 * i + 2] = sv.omega[i].z;
  }
  return 0;
}

// Vorticity confinement from a given state (for the closed-form checks):
// omega, eta = grad|omega| and the force f = eps_v (N x omega) per particle,
// with the full candidate lists of one grid build at x (no prediction), and
// v + dt f as the step applies it.
int32_t orc_vorticity_confinement(const orc_params* prm, int32_t n, const float* pos3, const float* vel3,
                                  const float* rho, float mass, float dt, float* omega3, float* eta3,
                                  float* force3, float* vnew3) {
  State<float> st;
  load_state<float>(st, n, pos3, vel3, rho, nullptr);
  Solver<float> sv(Params<float>::from(*prm));
  sv.grid.build(st.xs);
  sv.nbr.resize(n);
  std::vector<int> scratch;
  for (int i = 0; i < n; ++i) {
    sv.grid.query(st.xs[i], scratch);
    sv.nbr[i] = scratch;
  }
  sv.vorticity_omega(st, mass);
  std::vector<V3<float>> vnew(n);
  sv.vorticity_apply(st, mass, dt, vnew);
  for (int i = 0; i < n; ++i) {
    V3<float> eta;
    const V3<float> f = sv.vorticity_force(st, mass, i, &eta);
    const V3<float>* src[4] = {&sv.omega[i], &eta, &f, &vnew[i]};
    float* dst[4] = {omega3, eta3, force3, vnew3};
    for (int k = 0; k < 4; ++k) {
      dst[k][3 * i] = src[k]->x;
      dst[k][3 * i + 1] = src[k]->y;
      dst[k][3 * i + 2] = src[k]->z;
    }
  }
  return 0;
}

}  // extern "C"
