"""Akinci boundary path on the GPU (through the C ABI) against the oracle:
psi volumes, boundary world positions, boundary density and per-body coupling
force / torque after a PBF step, all bitwise."""
import os
import sys

import numpy as np
import pytest

import novaphy_amd as nv
from novaphy_amd import capi, scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _scene_with_bodies():
    """A small dam break with a ground plane, two spheres and a rotated box
    sitting in / on the fluid (coupling forces non-zero)."""
    sc = scenes.dam_break(16, 10, 16, (2.0, 1.0, 1.0), seed=11, steps=1)
    shapes = [
        O.make_shape(O.SHAPE_PLANE, -1, [0, 1, 0, 0.0]),
        O.make_shape(O.SHAPE_SPHERE, 0, [0.15, 0, 0, 0]),
        O.make_shape(O.SHAPE_SPHERE, 1, [0.1, 0, 0, 0]),
        O.make_shape(O.SHAPE_BOX, 2, [0.12, 0.08, 0.1, 0]),
    ]
    q = np.array([0.9238795, 0.0, 0.3826834, 0.0], np.float32)  # 45 deg about y
    poses = np.array([[0.35, 0.45, 0.40, 1, 0, 0, 0],
                      [0.60, 0.30, 0.55, 1, 0, 0, 0],
                      [0.45, 0.52, 0.70, *q]], np.float32)
    static = np.array([0, 0, 0], np.uint8)
    local, body = O.sample_boundaries(shapes, 0.05, 1.0)
    return sc, local, body, poses, static


def test_psi_volumes_bitexact(gpu):
    sc, local, body, poses, static = _scene_with_bodies()
    _say("case: psi volumes nb=", local.shape[0])
    c = capi.Context(sc.params, capacity=sc.n)
    c.set_particles(sc.positions, sc.velocities)
    c.set_boundary(local, body, None, nbodies=poses.shape[0])
    psi = c.compute_boundary_volumes(poses)
    wpos = O.boundary_world_positions(local, body, poses)
    opsi = O.boundary_volumes(wpos, sc.params.kernel_radius)
    np.testing.assert_array_equal(psi, opsi)
    gw = c.boundary_world_positions(poses, read_back=True)
    np.testing.assert_array_equal(gw, wpos)
    c.close()


def test_boundary_density_and_coupling_bitexact(gpu):
    sc, local, body, poses, static = _scene_with_bodies()
    h = sc.params.kernel_radius
    _say("case: boundary density + coupling n=", sc.n, "nb=", local.shape[0])
    c = capi.Context(sc.params, capacity=sc.n)
    c.set_particles(sc.positions, sc.velocities)
    c.set_boundary(local, body, None, nbodies=poses.shape[0])
    psi = c.compute_boundary_volumes(poses)
    # FluidWorld::step order (fluid_world.cpp:46-99)
    c.boundary_world_positions(poses)
    c.step(sc.dt, sc.gravity, sc.mass, 1)
    c.apply_boundary_density()
    f, t = c.coupling_forces(poses, static)
    gp, gv, gr, gl = c.get_particles()

    wpos = O.boundary_world_positions(local, body, poses)
    op, ov, orho, ol = O.pbf_run(sc.params, sc.positions, sc.velocities, sc.dt, sc.gravity, sc.mass, 1)
    orho_b = O.boundary_density(op, orho, wpos, psi, h, sc.params.rest_density)
    of, ot = O.coupling(op, orho_b, wpos, psi, body, poses, static, h, sc.params.rest_density, sc.mass)
    np.testing.assert_array_equal(gp, op)
    np.testing.assert_array_equal(gr, orho_b)
    assert np.abs(of).max() > 0, "the scene must produce coupling forces"
    np.testing.assert_array_equal(f, of)
    np.testing.assert_array_equal(t, ot)
    # FP64 recompute of the coupling for the record
    f64, t64 = O.coupling_f64(op, orho_b, wpos, psi, body, poses, static, h, sc.params.rest_density, sc.mass)
    _say("  coupling FP32 vs FP64: force max-rel %.2e" % (np.abs(f - f64).max() / np.abs(f64).max()))
    # static bodies receive nothing
    f2, t2 = c.coupling_forces(poses, np.array([1, 0, 1], np.uint8))
    assert np.all(f2[0] == 0) and np.all(f2[2] == 0)
    np.testing.assert_array_equal(f2[1], of[1])
    c.close()


@pytest.mark.parametrize("huge_levels", [6, 0])
def test_coupling_with_stacked_fluid_bitexact(gpu, huge_levels):
    """Coupling forces when runs of identical fluid particles touch the bodies
    (two stacks of 1200 and 400 particles on one point each, under the two
    spheres): the coupling walks the fluid candidates run by run and adds a
    run's term k times over, which must equal the oracle's one-by-one sums bit
    for bit, in a single context and in a batch.  huge_levels 0: every class
    spanning two or more blocks (the 1200-particle stack's among them) ranks
    in index space in the step's sort and in the coupling's fluid and
    boundary sorts."""
    sc, local, body, poses, static = _scene_with_bodies()
    pos = sc.positions.copy()
    pos[:1200] = np.array([0.35, 0.28, 0.40], np.float32)
    pos[1200:1600] = np.array([0.62, 0.18, 0.55], np.float32)
    vel = np.zeros_like(pos)
    h = sc.params.kernel_radius
    _say("case: coupling with stacked fluid n=", sc.n, "nb=", local.shape[0], "huge_levels", huge_levels)
    wpos = O.boundary_world_positions(local, body, poses)
    psi = O.boundary_volumes(wpos, h)
    op, ov, orho, ol = O.pbf_run(sc.params, pos, vel, sc.dt, sc.gravity, sc.mass, 1)
    orho_b = O.boundary_density(op, orho, wpos, psi, h, sc.params.rest_density)
    of, ot = O.coupling(op, orho_b, wpos, psi, body, poses, static, h, sc.params.rest_density, sc.mass)
    assert np.abs(of[0]).max() > 0 and np.abs(of[1]).max() > 0
    assert np.unique(op[:1200], axis=0).shape[0] == 1  # the stack stays one point

    c = capi.Context(sc.params, capacity=sc.n)
    c.set_rank_split(huge_levels)
    c.set_particles(pos, vel)
    c.set_boundary(local, body, psi, nbodies=poses.shape[0])
    c.boundary_world_positions(poses)
    c.step(sc.dt, sc.gravity, sc.mass, 1)
    c.apply_boundary_density()
    f, t = c.coupling_forces(poses, static)
    gp, _, gr, _ = c.get_particles()
    c.close()
    np.testing.assert_array_equal(gp, op)
    np.testing.assert_array_equal(gr, orho_b)
    np.testing.assert_array_equal(f, of)
    np.testing.assert_array_equal(t, ot)

    slots = ((sc.n + 255) // 256) * 256
    b = capi.Batch(slot_capacity=slots, cell_capacity=capi.scene_bytes(sc.params, 1), max_scenes=2)
    b.set_rank_split(huge_levels)
    sid = b.add(sc.params, sc.dt, sc.gravity, sc.mass, pos, vel)
    b.set_boundary(sid, local, body, psi, poses.shape[0], static)
    b.step(1)
    out = b.coupling(poses)
    bp, _, br, _ = b.get(sid)
    b.close()
    np.testing.assert_array_equal(bp, op)
    np.testing.assert_array_equal(br, orho_b)
    np.testing.assert_array_equal(out[:, :3], of)
    np.testing.assert_array_equal(out[:, 3:], ot)


def test_batch_persistent_passes_equal_alone(gpu):
    """A batch stepped with persistent neighbour passes (16 workgroups per
    pass, each looping over particle blocks of several scenes) and skin lists
    (list policy 1) matches each scene alone at the default launch shape,
    bitwise."""
    dims = [(6, 9, 5), (12, 7, 10), (9, 9, 9), (20, 6, 8)]
    scs = [scenes.batched_scene(nx, ny, nz, 0.4, seed=200 + k, steps=6) for k, (nx, ny, nz) in enumerate(dims)]
    _say("case: persistent batch of", len(scs), "scenes, sizes", [s.n for s in scs])
    slots = sum(((s.n + 255) // 256) * 256 for s in scs)
    cells = sum(capi.scene_bytes(s.params, 1) for s in scs)  # generous upper bound on cells
    b = capi.Batch(slot_capacity=slots, cell_capacity=cells, max_scenes=8)
    b.set_tuning(128, 1, 0.1, 16)
    b.set_list_policy(1)
    ids = [b.add(s.params, s.dt, s.gravity, s.mass, s.positions, s.velocities) for s in scs]
    b.step(6)
    for s, sid in zip(scs, ids):
        bp, bv, br, _ = b.get(sid)
        c = capi.Context(s.params, capacity=s.n)
        c.set_particles(s.positions, s.velocities)
        c.step(s.dt, s.gravity, s.mass, 6)
        ap, av, ar, _ = c.get_particles()
        c.close()
        np.testing.assert_array_equal(bp, ap)
        np.testing.assert_array_equal(bv, av)
        np.testing.assert_array_equal(br, ar)
    b.close()


def test_batched_scenes_equal_alone(gpu):
    """Each scene advanced inside a batch matches the same scene alone,
    bitwise (per-scene arithmetic is independent of the batch)."""
    dims = [(6, 9, 5), (12, 7, 10), (9, 9, 9), (20, 6, 8)]
    scs = [scenes.batched_scene(nx, ny, nz, 0.4, seed=100 + k, steps=10) for k, (nx, ny, nz) in enumerate(dims)]
    _say("case: batch of", len(scs), "scenes, sizes", [s.n for s in scs])
    slots = sum(((s.n + 255) // 256) * 256 for s in scs)
    cells = sum(capi.scene_bytes(s.params, 1) for s in scs)  # generous upper bound on cells
    b = capi.Batch(slot_capacity=slots, cell_capacity=cells, max_scenes=8)
    ids = [b.add(s.params, s.dt, s.gravity, s.mass, s.positions, s.velocities) for s in scs]
    b.step(10)
    for s, sid in zip(scs, ids):
        bp, bv, br, bl = b.get(sid)
        c = capi.Context(s.params, capacity=s.n)
        c.set_particles(s.positions, s.velocities)
        c.step(s.dt, s.gravity, s.mass, 10)
        ap, av, ar, al = c.get_particles()
        c.close()
        np.testing.assert_array_equal(bp, ap)
        np.testing.assert_array_equal(bv, av)
        np.testing.assert_array_equal(br, ar)
    # remove one scene, add another, step again: survivors continue unchanged
    b.remove(ids[1])
    extra = scenes.batched_scene(7, 7, 7, 0.3, seed=999, steps=5)
    eid = b.add(extra.params, extra.dt, extra.gravity, extra.mass, extra.positions, extra.velocities)
    b.step(5)
    ep, _, _, _ = b.get(eid)
    c = capi.Context(extra.params, capacity=extra.n)
    c.set_particles(extra.positions, extra.velocities)
    c.step(extra.dt, extra.gravity, extra.mass, 5)
    np.testing.assert_array_equal(ep, c.get_particles()[0])
    c.close()
    s0 = scs[0]
    c = capi.Context(s0.params, capacity=s0.n)
    c.set_particles(s0.positions, s0.velocities)
    c.step(s0.dt, s0.gravity, s0.mass, 15)
    np.testing.assert_array_equal(b.get(ids[0])[0], c.get_particles()[0])
    c.close()
    b.close()


def test_device_boundary_sampling_matches_host(gpu):
    """npx_sample_boundaries against the host samplers (boundary.cpp:15-150):
    the same particles in the same order and bodies, every sample bit for
    bit (the sphere samples' trigonometry is the host libm's float
    functions, as in the reference's sampler)."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from test_capi_cpu import _shape_model
    m = _shape_model()
    _say("case: device boundary sampling vs host, shapes", m.num_shapes)
    shapes, n = capi.shapes_from_model(m)
    local, body = capi.sample_boundaries(shapes, n, 0.05, 0.7)
    host = nv.sample_model_boundaries(m, 0.05, 0.7)
    hl = np.array([np.array(b.local_position, np.float32) for b in host], np.float32)
    hb = np.array([b.body_index for b in host], np.int32)
    assert local.shape == hl.shape
    np.testing.assert_array_equal(body, hb)
    kinds = []
    for sh in m.shapes:
        k = str(sh.type).split(".")[-1]
        cnt = len(nv.sample_model_boundaries(_single(sh), 0.05, 0.7))
        kinds += [k] * cnt
    kinds = np.array(kinds)
    sph = kinds == "Sphere"
    assert sph.sum() > 100 and (~sph).sum() > 100
    _say("  sphere samples differing: %d of %d" % (int(np.any(local[sph] != hl[sph], axis=1).sum()), int(sph.sum())))
    np.testing.assert_array_equal(local, hl)


def _single(shape):
    """A model holding only `shape` (and the bodies it refers to)."""
    b = nv.ModelBuilder()
    for _ in range(max(shape.body_index + 1, 0)):
        b.add_body(nv.RigidBody.from_sphere(1.0, 0.1))
    b.add_shape(shape)
    return b.build()


def test_batch_mixed_iterations_and_stacks(gpu):
    """A batch whose scenes differ in solver iterations (0, 2, 4) and hold
    stacks of identical particles: every scene matches itself alone bitwise,
    and the scene with no solver iteration matches the oracle."""
    import dataclasses
    rng = np.random.default_rng(41)
    base = scenes.batched_scene(9, 8, 7, 0.4, seed=301, steps=6)
    stacked = scenes.batched_scene(10, 6, 9, 0.4, seed=302, steps=6)
    pos = stacked.positions.copy()
    vel = stacked.velocities.copy()
    pick = rng.choice(pos.shape[0], 150, replace=False)
    pos[pick] = pos[pick[0]]  # 150 particles on one point with one velocity
    vel[pick] = vel[pick[0]]
    stacked = dataclasses.replace(stacked, positions=pos, velocities=vel)
    zero = dataclasses.replace(base, params=dataclasses.replace(base.params, solver_iterations=0))
    two = scenes.batched_scene(8, 9, 6, 0.4, seed=303, steps=6)
    two = dataclasses.replace(two, params=dataclasses.replace(two.params, solver_iterations=2))
    scs = [base, stacked, zero, two]
    _say("case: batch with 0 / 2 / 4 iterations and a stack, sizes", [s.n for s in scs])
    slots = sum(((s.n + 255) // 256) * 256 for s in scs)
    cells = sum(capi.scene_bytes(s.params, 1) for s in scs)
    b = capi.Batch(slot_capacity=slots, cell_capacity=cells, max_scenes=8)
    ids = [b.add(s.params, s.dt, s.gravity, s.mass, s.positions, s.velocities) for s in scs]
    b.step(6)
    for s, sid in zip(scs, ids):
        bp, bv, br, bl = b.get(sid)
        c = capi.Context(s.params, capacity=s.n)
        c.set_particles(s.positions, s.velocities)
        c.step(s.dt, s.gravity, s.mass, 6)
        ap, av, ar, al = c.get_particles()
        c.close()
        np.testing.assert_array_equal(bp, ap)
        np.testing.assert_array_equal(bv, av)
        np.testing.assert_array_equal(br, ar)
        np.testing.assert_array_equal(bl, al)
    op, ov, _, _ = O.pbf_run(zero.params, zero.positions, zero.velocities, zero.dt, zero.gravity, zero.mass, 6)
    zp, zv, _, _ = b.get(ids[2])
    np.testing.assert_array_equal(zp, op)
    np.testing.assert_array_equal(zv, ov)
    b.close()


def test_batch_crowded_scenes_equal_alone(gpu):
    """Crowded wall sheets in a batch beside an ordinary scene, with 2 and 4
    solver iterations: the dense list spans scenes, and the dense kernels
    skip scenes past their iteration count.  Every scene matches itself
    alone bitwise."""
    rng = np.random.default_rng(61)

    def sheet(n, iters, vort):
        pos = np.empty((n, 3), np.float32)
        pos[:, 0] = np.float32(0.025)
        pos[:, 1:] = (0.2 + rng.uniform(0.0, 0.3, (n, 2))).astype(np.float32)
        vel = rng.uniform(-0.3, 0.3, (n, 3)).astype(np.float32)
        prm = scenes.PBFParams(corr_k=0.1, corr_delta_q=0.2, vorticity_epsilon=vort, solver_iterations=iters,
                               use_domain_bounds=True, domain_lower=(0.0, 0.0, 0.0), domain_upper=(1.0, 1.0, 1.0))
        return scenes.Scene("sheet", pos, vel, prm, float(np.float32(1.0 / 120.0)), (0.0, -9.81, 0.0),
                            prm.particle_mass(0.025), 3, 61)

    plain = scenes.batched_scene(9, 8, 7, 0.4, seed=311, steps=4)
    scs = [sheet(2500, 2, 0.0), plain, sheet(2000, 4, 1e-3)]
    _say("case: batch with crowded sheets, sizes", [s.n for s in scs])
    slots = sum(((s.n + 255) // 256) * 256 for s in scs)
    cells = sum(capi.scene_bytes(s.params, 1) for s in scs)
    b = capi.Batch(slot_capacity=slots, cell_capacity=cells, max_scenes=4)
    ids = [b.add(s.params, s.dt, s.gravity, s.mass, s.positions, s.velocities) for s in scs]
    b.step(3)
    for s, sid in zip(scs, ids):
        bp, bv, br, bl = b.get(sid)
        c = capi.Context(s.params, capacity=s.n)
        c.set_particles(s.positions, s.velocities)
        c.step(s.dt, s.gravity, s.mass, 3)
        ap, av, ar, al = c.get_particles()
        c.close()
        np.testing.assert_array_equal(bp, ap)
        np.testing.assert_array_equal(bv, av)
        np.testing.assert_array_equal(br, ar)
        np.testing.assert_array_equal(bl, al)
    b.close()
