// Drives the oracle's C entry points (oracle/pbf_oracle.cpp, rigid_oracle.cpp)
// on small inputs for an AddressSanitizer / UBSan build: PBF steps with the
// configured s_corr and vorticity, grid diagnostics and queries, boundary
// sampling, volumes, density and coupling, and rigid steps with contacts.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

extern "C" {
struct orc_params {
  float rest_density, kernel_radius, particle_radius;
  int32_t solver_iterations;
  float epsilon, xsph_viscosity, vorticity_epsilon, corr_delta_q, corr_k;
  int32_t corr_n, use_domain_bounds;
  float domain_lower[3], domain_upper[3];
};
struct orc_shape {
  int32_t type, body;
  float a[4], local_pos[3], local_rot[4];
};
struct orc_rbody {
  float mass, inertia[9];
};
struct orc_rshape {
  int32_t type, body;
  float half[3], radius, normal[3], offset, local_pos[3], local_rot[4], friction, restitution;
};
struct orc_rsettings {
  int32_t velocity_iterations;
  float baumgarte, slop;
  int32_t warm_starting;
};
struct orc_contact {
  float position[3], normal[3], penetration;
  int32_t body_a, body_b;
  float friction, restitution, acc_n, acc_t1, acc_t2;
};
int32_t orc_generate_fluid_block(const float*, const float*, float, float*);
int32_t orc_pbf_run_f32(const orc_params*, int32_t, float*, float*, float*, float*, float, const float*, float,
                        int32_t);
int64_t orc_grid_debug(int32_t, const float*, float, float, uint64_t*, int32_t*, int32_t*, int32_t*, int64_t);
int64_t orc_grid_all_pairs(int32_t, const float*, float, float, int32_t*, int64_t);
int32_t orc_sample_boundaries(int32_t, const orc_shape*, float, float, float*, int32_t*);
int32_t orc_boundary_world_positions(int32_t, const float*, const int32_t*, int32_t, const float*, float*);
int32_t orc_boundary_volumes(int32_t, const float*, float, float*);
int32_t orc_boundary_density(int32_t, const float*, float*, int32_t, const float*, const float*, float, float);
int32_t orc_coupling(int32_t, const float*, const float*, int32_t, const float*, const float*, const int32_t*, int32_t,
                     const float*, const uint8_t*, float, float, float, float*, float*, float*);
int32_t orc_rigid_step(const orc_rbody*, int32_t, const orc_rshape*, int32_t, const orc_rsettings*, const float*,
                       float, float*, float*, float*, float*, float*, orc_contact*, int32_t, int32_t*, int32_t*);
}

int main() {
  const float lo[3] = {0.025f, 0.025f, 0.025f}, hi[3] = {0.5f, 0.4f, 0.5f};
  const int n = orc_generate_fluid_block(lo, hi, 0.05f, nullptr);
  std::vector<float> pos(3 * n), vel(3 * n, 0.0f), rho(n), lam(n);
  orc_generate_fluid_block(lo, hi, 0.05f, pos.data());
  orc_params p{1000.0f, 0.1f, 0.025f, 4, 100.0f, 0.01f, 1e-4f, 0.2f, 0.1f, 4, 1, {0, 0, 0}, {0.6f, 0.6f, 0.6f}};
  const float g[3] = {0.0f, -9.81f, 0.0f};
  orc_pbf_run_f32(&p, n, pos.data(), vel.data(), rho.data(), lam.data(), 1.0f / 120.0f, g, 0.125f, 12);
  std::vector<uint64_t> keys(n);
  std::vector<int32_t> cc(n), nc(n);
  const int64_t tot = orc_grid_debug(n, pos.data(), 0.1f, 0.1f, keys.data(), cc.data(), nc.data(), nullptr, 0);
  std::vector<int32_t> lists(tot);
  orc_grid_debug(n, pos.data(), 0.1f, 0.1f, keys.data(), cc.data(), nc.data(), lists.data(), tot);
  const int64_t np = orc_grid_all_pairs(n, pos.data(), 0.1f, 0.1f, nullptr, 0);
  std::vector<int32_t> pairs(2 * np);
  orc_grid_all_pairs(n, pos.data(), 0.1f, 0.1f, pairs.data(), np);
  orc_shape sh[3] = {{2, -1, {0, 1, 0, 0}, {0, 0, 0}, {1, 0, 0, 0}},
                     {1, 0, {0.1f, 0, 0, 0}, {0, 0, 0}, {1, 0, 0, 0}},
                     {0, 1, {0.08f, 0.05f, 0.1f, 0}, {0.01f, 0, 0}, {0.98f, 0.1f, 0.1f, 0.1f}}};
  const int nb = orc_sample_boundaries(3, sh, 0.05f, 0.6f, nullptr, nullptr);
  std::vector<float> local(3 * nb), wpos(3 * nb), psi(nb);
  std::vector<int32_t> body(nb);
  orc_sample_boundaries(3, sh, 0.05f, 0.6f, local.data(), body.data());
  const float pose[14] = {0.3f, 0.35f, 0.3f, 1, 0, 0, 0, 0.2f, 0.3f, 0.4f, 0.9238795f, 0, 0.3826834f, 0};
  orc_boundary_world_positions(nb, local.data(), body.data(), 2, pose, wpos.data());
  orc_boundary_volumes(nb, wpos.data(), 0.1f, psi.data());
  orc_boundary_density(n, pos.data(), rho.data(), nb, wpos.data(), psi.data(), 0.1f, 1000.0f);
  const uint8_t st[2] = {0, 0};
  float f[6], t[6];
  std::vector<float> pp(3 * nb);
  orc_coupling(n, pos.data(), rho.data(), nb, wpos.data(), psi.data(), body.data(), 2, pose, st, 0.1f, 1000.0f,
               0.125f, f, t, pp.data());
  // rigid: boxes and spheres falling on a plane and on each other
  const int nbod = 6;
  std::vector<orc_rbody> bodies(nbod);
  std::vector<orc_rshape> shapes(nbod + 1);
  shapes[0] = orc_rshape{2, -1, {0, 0, 0}, 0, {0, 1, 0}, 0, {0, 0, 0}, {1, 0, 0, 0}, 0.5f, 0.0f};
  std::vector<float> pose7(7 * nbod), lin(3 * nbod, 0.0f), ang(3 * nbod, 0.0f), frc(3 * nbod, 0.0f),
      trq(3 * nbod, 0.0f);
  for (int k = 0; k < nbod; ++k) {
    const bool box = k % 2 == 0;
    bodies[k] = orc_rbody{1.0f + k, {0.01f, 0, 0, 0, 0.012f, 0, 0, 0, 0.014f}};
    shapes[k + 1] = orc_rshape{box ? 0 : 1, k, {0.1f, 0.08f, 0.12f}, 0.1f, {0, 1, 0}, 0, {0, 0, 0},
                               {1, 0, 0, 0}, 0.5f, 0.3f};
    const float pk[7] = {0.05f * k, 0.15f + 0.22f * k, 0.03f * k, 0.98f, 0.1f * k, 0.05f, 0.1f};
    const float nq = std::sqrt(pk[3] * pk[3] + pk[4] * pk[4] + pk[5] * pk[5] + pk[6] * pk[6]);
    for (int d = 0; d < 7; ++d) pose7[7 * k + d] = d < 3 ? pk[d] : pk[d] / nq;
  }
  orc_rsettings rs{30, 0.3f, 0.005f, 1};
  std::vector<orc_contact> cont(64);
  int32_t ncont = 0, npair = 0, total = 0;
  for (int s = 0; s < 240; ++s) {
    orc_rigid_step(bodies.data(), nbod, shapes.data(), nbod + 1, &rs, g, 1.0f / 120.0f, pose7.data(), lin.data(),
                   ang.data(), frc.data(), trq.data(), cont.data(), (int32_t)cont.size(), &ncont, &npair);
    total += ncont;
  }
  std::printf("oracle driver ok: n=%d candidates=%lld pairs=%lld boundary=%d contacts=%d\n", n, (long long)tot,
              (long long)np, nb, total);
  return total > 0 ? 0 : 1;
}
