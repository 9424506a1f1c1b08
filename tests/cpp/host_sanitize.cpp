// Drives the C++ host's CPU paths (novaphy_amd/csrc/host: rigid pipeline,
// sweep and prune, narrowphase, boundary samplers, fluid blocks, performance
// monitor and its trace) for an AddressSanitizer / UBSan build.  No GPU work:
// the libnpx_hip calls made here are the ones that fail before any device
// access.
#include <cstdio>
#include <string>

#include "npx_host.hpp"

using namespace novaphy;

int main(int argc, char** argv) {
  const std::string trace = argc > 1 ? argv[1] : "host_sanitize_trace.json";
  ModelBuilder b;
  b.add_ground_plane(0.0f, 0.6f, 0.1f);
  for (int k = 0; k < 10; ++k) {
    const Vec3f p(0.05f * (k % 3), 0.2f + 0.3f * k, 0.04f * (k % 2));
    const Quatf q = Quatf(0.97f, 0.1f * (k % 3), 0.15f, -0.05f * k).normalized();
    const Transform local(Vec3f(0.01f, -0.01f, 0.0f), Quatf(0.99f, 0.05f, 0.0f, 0.1f));
    if (k % 2 == 0) {
      const Vec3f he(0.1f + 0.01f * k, 0.08f, 0.12f);
      const int i = b.add_body(k == 4 ? RigidBody::make_static() : RigidBody::from_box(1.0f + k, he), Transform(p, q));
      b.add_shape(CollisionShape::make_box(he, i, local));
    } else {
      const int i = b.add_body(RigidBody::from_sphere(0.5f + k, 0.1f), Transform(p, q));
      b.add_shape(CollisionShape::make_sphere(0.1f, i, local));
      b.add_shape(CollisionShape::make_sphere(0.05f, i));  // two shapes on one body
    }
  }
  const Model model = b.build();
  World w(model);
  w.performance_monitor().enabled = true;
  w.performance_monitor().trace_enabled = true;
  size_t contacts = 0;
  for (int s = 0; s < 300; ++s) {
    w.apply_force(s % 10, Vec3f(0.1f, 0.0f, -0.1f));
    w.apply_torque(s % 10, Vec3f(0.0f, 0.01f, 0.0f));
    w.apply_force(99, Vec3f(1.0f, 0.0f, 0.0f));  // ignored, as in the reference
    w.step(1.0f / 120.0f);
    contacts += w.contacts().size();
  }
  w.performance_monitor().write_trace_json(trace);
  auto bp = sample_model_boundaries(model, 0.05f, 0.5f);
  FluidBlockDef blk;
  blk.lower = Vec3f(0.0f, 0.0f, 0.0f);
  blk.upper = Vec3f(0.3f, 0.2f, 0.25f);
  blk.particle_spacing = 0.05f;
  auto fluid = generate_fluid_block(blk);
  npx_pbf_params bad{};
  npx_ctx* ctx = nullptr;
  const int rc = npx_create(&bad, 16, 0, &ctx);
  std::printf("host driver ok: contacts=%zu boundary=%zu fluid=%zu create(bad)=%d\n", contacts, bp.size(),
              fluid.size(), rc);
  return (contacts > 0 && rc < 0) ? 0 : 1;
}
