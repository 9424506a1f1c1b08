"""Host-side API of novaphy_amd._core that runs without a GPU: the reference's
Python tests for kernels, blocks, settings, particle state, boundary sampling
and the rigid World, plus host sampling == oracle sampling bitwise."""
import math

import numpy as np
import numpy.testing as npt
import pytest

import novaphy_amd as novaphy
from oracle import oracle as O


def test_sph_kernels_reference_properties():  # test_fluid.py:11-91
    h = 0.1
    assert novaphy.SPHKernels.poly6(0.0, h) > 0
    assert abs(novaphy.SPHKernels.poly6(h * h, h)) < 1e-10
    assert novaphy.SPHKernels.poly6(h * h * 1.1, h) == 0.0
    n = 1000
    dr = h / n
    integral = sum(novaphy.SPHKernels.poly6(((i + 0.5) * dr) ** 2, h) * 4 * math.pi * ((i + 0.5) * dr) ** 2 * dr
                   for i in range(n))
    npt.assert_allclose(integral, 1.0, atol=0.02)
    assert novaphy.SPHKernels.spiky(0.0, h) > 0
    assert novaphy.SPHKernels.spiky(h, h) == 0.0
    npt.assert_allclose(novaphy.SPHKernels.spiky_grad(np.array([h, 0, 0], np.float32), h), [0, 0, 0], atol=1e-7)
    assert novaphy.SPHKernels.spiky_grad(np.array([0.05, 0, 0], np.float32), h)[0] < 0
    assert novaphy.SPHKernels.cubic_spline(0.0, h) > 0
    assert novaphy.SPHKernels.cubic_spline(h, h) == 0.0
    for r2 in (0.0, 0.0013, 0.0049, 0.0099):
        assert novaphy.SPHKernels.poly6(r2, h) == O.poly6(r2, h)


def test_fluid_block_reference_properties():  # test_fluid.py:96-144
    block = novaphy.FluidBlockDef()
    npt.assert_allclose(block.particle_spacing, 0.02, atol=1e-6)
    npt.assert_allclose(block.rest_density, 1000.0, atol=1e-3)
    npt.assert_allclose(block.lower, [0, 0, 0])
    npt.assert_allclose(block.upper, [0, 0, 0])
    block.lower = np.array([0, 0, 0], dtype=np.float32)
    block.upper = np.array([0.1, 0.1, 0.1], dtype=np.float32)
    block.particle_spacing = 0.05
    assert len(novaphy.generate_fluid_block(block)) == 27
    block.upper = np.array([0.2, 0.0, 0.0], dtype=np.float32)
    block.particle_spacing = 0.1
    p = novaphy.generate_fluid_block(block)
    assert len(p) == 3
    for a, b in zip(p, p[1:]):
        npt.assert_allclose(np.linalg.norm(b - a), 0.1, atol=1e-5)
    block.lower = np.array([0.025] * 3, np.float32)
    block.upper = np.array([0.975] * 3, np.float32)
    block.particle_spacing = 0.05
    np.testing.assert_array_equal(np.array(novaphy.generate_fluid_block(block)),
                                  O.generate_fluid_block([0.025] * 3, [0.975] * 3, 0.05))


def test_particle_state_reference_properties():  # test_fluid.py:149-183
    st = novaphy.ParticleState()
    st.init([np.array(v, np.float32) for v in ([0, 0, 0], [1, 0, 0], [0, 1, 0])])
    assert st.num_particles == 3 and len(st.positions) == 3 and len(st.densities) == 3
    st.init([np.array([0, 0, 0], np.float32)], np.array([1.0, 2.0, 3.0], np.float32))
    npt.assert_allclose(st.velocities[0], [1, 2, 3], atol=1e-6)
    st.clear()
    assert st.num_particles == 0
    st.init_arrays(np.zeros((5, 3), np.float32), np.ones((5, 3), np.float32))
    assert st.positions_numpy().shape == (5, 3) and st.velocities_numpy()[4, 2] == 1.0


def test_pbf_settings_reference_properties():  # test_pbf.py:11-27
    s = novaphy.PBFSettings()
    npt.assert_allclose(s.rest_density, 1000.0, atol=1e-3)
    npt.assert_allclose(s.kernel_radius, 0.1, atol=1e-6)
    assert s.solver_iterations == 4
    npt.assert_allclose(s.particle_mass(0.05), 0.125, rtol=1e-5)
    # C++-only fields of the reference are bound here (superset)
    assert s.vorticity_epsilon == 0.0 and s.corr_n == 4
    npt.assert_allclose([s.corr_delta_q, s.corr_k], [0.3, 0.001], rtol=1e-6)
    assert s.use_domain_bounds is False
    npt.assert_allclose(s.domain_lower, [-10, 0, -10])


def _model(shapes_fn):
    b = novaphy.ModelBuilder()
    shapes_fn(b)
    return b.build()


def _oracle_shape(s):
    t = s.local_transform
    rot = t.rotation  # [x, y, z, w]
    lp = list(t.position)
    lr = [rot[3], rot[0], rot[1], rot[2]]
    if s.type == novaphy.ShapeType.Box:
        return O.make_shape(O.SHAPE_BOX, s.body_index, list(s.half_extents) + [0], lp, lr)
    if s.type == novaphy.ShapeType.Sphere:
        return O.make_shape(O.SHAPE_SPHERE, s.body_index, [s.radius, 0, 0, 0], lp, lr)
    return O.make_shape(O.SHAPE_PLANE, s.body_index, list(s.normal) + [s.offset], lp, lr)


def test_boundary_sampling_reference_properties_and_oracle():  # test_coupling.py:10-78
    half = np.array([0.5, 0.5, 0.5], np.float32)

    def box(b):
        i = b.add_body(novaphy.RigidBody.from_box(1.0, half))
        b.add_shape(novaphy.CollisionShape.make_box(half, i))

    def sphere(b):
        i = b.add_body(novaphy.RigidBody.from_sphere(1.0, 0.5))
        b.add_shape(novaphy.CollisionShape.make_sphere(0.5, i))

    def rotated(b):
        b.add_ground_plane(y=0.1)
        i = b.add_body(novaphy.RigidBody.from_box(2.0, half))
        loc = novaphy.Transform(np.array([0.1, 0.0, -0.2], np.float32), np.array([0.0, 0.3826834, 0.0, 0.9238795]))
        b.add_shape(novaphy.CollisionShape.make_box(np.array([0.2, 0.1, 0.15], np.float32), i, loc))
        j = b.add_body(novaphy.RigidBody.from_sphere(1.0, 0.12))
        b.add_shape(novaphy.CollisionShape.make_sphere(0.12, j))

    m = _model(box)
    p = novaphy.sample_model_boundaries(m, 0.25)
    assert len(p) > 0 and all(x.body_index == 0 for x in p)
    m = _model(sphere)
    p = novaphy.sample_model_boundaries(m, 0.1)
    assert len(p) > 10
    for x in p:
        npt.assert_allclose(np.linalg.norm(np.array(x.local_position)), 0.5, atol=0.01)
    m = _model(lambda b: b.add_ground_plane(y=0.0))
    p = novaphy.sample_model_boundaries(m, 0.2, plane_extent=0.5)
    assert len(p) > 0 and all(x.body_index == -1 for x in p)
    for x in p:
        npt.assert_allclose(x.local_position[1], 0.0, atol=1e-5)
    assert len(novaphy.sample_model_boundaries(novaphy.ModelBuilder().build(), 0.1)) == 0
    bp = novaphy.BoundaryParticle()
    npt.assert_allclose(bp.local_position, [0, 0, 0])
    assert bp.body_index == -1 and bp.volume == 0.0
    # the product's sampler equals the oracle's restatement bit for bit
    for fn, s, ext in ((box, 0.25, 1.0), (sphere, 0.1, 1.0), (rotated, 0.05, 0.7)):
        m = _model(fn)
        p = novaphy.sample_model_boundaries(m, s, plane_extent=ext)
        local, body = O.sample_boundaries([_oracle_shape(x) for x in m.shapes], s, ext)
        np.testing.assert_array_equal(np.array([x.local_position for x in p], np.float32), local)
        np.testing.assert_array_equal(np.array([x.body_index for x in p], np.int32), body)


def test_world_rigid_pipeline_cpu():
    b = novaphy.ModelBuilder()
    b.add_ground_plane()
    half = np.array([0.5, 0.5, 0.5], np.float32)
    i = b.add_body(novaphy.RigidBody.from_box(1.0, half),
                   novaphy.Transform.from_translation(np.array([0.0, 5.0, 0.0], np.float32)))
    b.add_shape(novaphy.CollisionShape.make_box(half, i))
    j = b.add_body(novaphy.RigidBody.from_sphere(1.0, 0.3),
                   novaphy.Transform.from_translation(np.array([3.0, 2.0, 0.0], np.float32)))
    b.add_shape(novaphy.CollisionShape.make_sphere(0.3, j))
    w = novaphy.World(b.build())
    for _ in range(240):
        w.step(1.0 / 60.0)
    y = [t.position[1] for t in w.state.transforms]
    assert 0.45 < y[0] < 0.55     # box at rest on the plane
    assert 0.25 < y[1] < 0.35     # sphere at rest on the plane


def test_fluid_world_without_fluid_runs_on_host():  # test_pbf.py:139-170
    w = novaphy.FluidWorld(novaphy.ModelBuilder().build())
    assert w.num_particles == 0
    w.step(1.0 / 60.0)
    b = novaphy.ModelBuilder()
    b.add_ground_plane()
    half = np.array([0.5, 0.5, 0.5], np.float32)
    i = b.add_body(novaphy.RigidBody.from_box(1.0, half),
                   novaphy.Transform.from_translation(np.array([0.0, 5.0, 0.0], np.float32)))
    b.add_shape(novaphy.CollisionShape.make_box(half, i))
    w = novaphy.FluidWorld(b.build())
    y0 = w.state.transforms[0].position[1]
    for _ in range(10):
        w.step(1.0 / 60.0)
    assert w.state.transforms[0].position[1] < y0
    assert w.num_boundary_particles > 0


def test_world_performance_monitor_cpu(tmp_path):
    b = novaphy.ModelBuilder()
    b.add_ground_plane()
    i = b.add_body(novaphy.RigidBody.from_sphere(1.0, 0.2),
                   novaphy.Transform.from_translation(np.array([0.0, 0.19, 0.0], np.float32)))
    b.add_shape(novaphy.CollisionShape.make_sphere(0.2, i))
    w = novaphy.World(b.build())
    mon = w.performance_monitor
    assert mon.enabled is False
    w.step(1 / 120)
    assert mon.phase_stats() == [] and mon.frame_count == 0
    mon.enabled = True
    mon.trace_enabled = True
    w.step(1 / 120)
    names = [p.name for p in mon.phase_stats()]
    assert "world.total" in names and "world.broadphase.sap" in names
    assert mon.frame_count == 1
    mon.reset()
    assert mon.phase_stats() == [] and mon.last_frame_metrics() == [] and mon.frame_count == 0
    out = tmp_path / "t.json"
    mon.write_trace_json(str(out))
    import json
    assert json.loads(out.read_text())["traceEvents"] == []


def _coupling_worlds():
    """Three rigid worlds of different body counts: spheres and boxes over a
    ground plane, a static body in the second."""
    rng = np.random.default_rng(7)
    worlds = []
    for nb, with_static in ((2, False), (3, True), (1, False)):
        b = novaphy.ModelBuilder()
        b.add_ground_plane()
        for k in range(nb):
            pos = np.array([rng.uniform(-1, 1), rng.uniform(0.5, 2.0), rng.uniform(-1, 1)], np.float32)
            body = novaphy.RigidBody.make_static() if (with_static and k == 1) else \
                novaphy.RigidBody.from_sphere(float(rng.uniform(0.5, 2.0)), 0.2)
            i = b.add_body(body, novaphy.Transform.from_translation(pos))
            b.add_shape(novaphy.CollisionShape.make_sphere(0.2, i))
        worlds.append(novaphy.World(b.build()))
    return worlds


def _state(worlds):
    return [np.array([np.r_[t.position, t.rotation] for t in w.state.transforms] +
                     [np.array(v) for v in w.state.linear_velocities] +
                     [np.array(v) for v in w.state.angular_velocities], dtype=object) for w in worlds]


def test_batched_coupling_round_equals_per_body_loop():
    """gather_poses7 / apply_forces_and_step (one host call each per served
    coupling round) equal the per-body loop of FluidWorld::step's rules:
    poses world by world as (position, w, x, y, z); static bodies get no force;
    a body whose force and torque both have a zero float32 squared norm
    (x*x + (y*y + z*z)), tiny values whose squares underflow included, gets
    none; then each world steps with its own dt."""
    a, b = _coupling_worlds(), _coupling_worlds()
    dts = [1.0 / 60.0, 1.0 / 120.0, 1.0 / 90.0]
    rng = np.random.default_rng(3)
    for rnd in range(4):
        poses = novaphy.gather_poses7(a)
        want = np.concatenate([np.array([[*t.position, t.rotation[3], *t.rotation[:3]] for t in w.state.transforms],
                                        np.float32) for w in b])
        npt.assert_array_equal(poses, want)
        out = rng.normal(0.0, 5.0, (poses.shape[0], 6)).astype(np.float32)
        out[1] = 0.0                      # zero force and torque
        out[3, :3] = 1e-30                # squares underflow: skipped
        out[3, 3:] = 0.0
        out[4, :3] = 0.0                  # torque only: applied
        novaphy.apply_forces_and_step(a, out, dts)
        off = 0
        for w, dt in zip(b, dts):
            for i, body in enumerate(w.model.bodies):
                f, t = out[off + i, :3], out[off + i, 3:]
                sq = lambda v: v[0] * v[0] + (v[1] * v[1] + v[2] * v[2])
                if not body.is_static() and not (sq(f) == 0.0 and sq(t) == 0.0):
                    w.apply_force(i, f)
                    w.apply_torque(i, t)
            w.step(dt)
            off += len(w.model.bodies)
        for sa, sb in zip(_state(a), _state(b)):
            for x, y in zip(sa, sb):
                npt.assert_array_equal(np.asarray(x, np.float32), np.asarray(y, np.float32))
    with pytest.raises(ValueError):
        novaphy.apply_forces_and_step(a, np.zeros((2, 6), np.float32), dts)
    with pytest.raises(ValueError):
        novaphy.apply_forces_and_step(a, np.zeros((poses.shape[0], 6), np.float32), dts[:2])
