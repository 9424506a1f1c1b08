"""The workloads of the benchmark configurations on the GPU, checked against
the oracle and against single-scene runs:

* coupled FluidWorld rollouts with boxes and spheres, 50 steps, against the
  oracle's fluid step, boundary density, coupling and rigid pipeline
  (fluid_world.cpp:46-174, world.cpp:30-182), bitwise;
* the 500k-particle + 64-body coupling workload, one full FluidWorld step;
* the 1024-scene batch, a seeded sample against single contexts and the
  two smallest scenes against the oracle;
* served requests over the full size range (4k-256k particles, 0-8 spheres),
  a sample alone against in the mix;
* the 8000-particle dam break with the configured s_corr over 200 steps
  against the oracle's committed rollout (tests/golden);
* query_all_pairs on a real particle set against the oracle.
"""
import hashlib
import json
import math
import os

import numpy as np
import pytest

import novaphy_amd as nv
from novaphy_amd import capi, scenes, serve
from oracle import oracle as O
from rigid_models import boundary_arrays, oracle_rigid, pose7, world_state

pytestmark = pytest.mark.gpu
F32 = np.float32
HERE = os.path.dirname(os.path.abspath(__file__))


def _say(*a):
    print(*a, flush=True)


def _static_mask(model):
    return np.array([b.is_static() for b in model.bodies], np.uint8)


def _oracle_run(world, sc, psi):
    local, body, _ = boundary_arrays(world)
    model = world.model
    ow = oracle_rigid(model, model.initial_transforms, world.solver_settings)
    return O.FluidWorldRun(sc.params, sc.positions, sc.velocities, world.particle_mass, local, body, psi, ow,
                           _static_mask(model))


def _initial_psi(world, h):
    local, body, _ = boundary_arrays(world)
    poses = np.array([pose7(t) for t in world.model.initial_transforms], F32).reshape(-1, 7)
    return O.boundary_volumes(O.boundary_world_positions(local, body, poses), h)


def _compare(world, fw, k):
    st = world.fluid_state
    np.testing.assert_array_equal(st.positions_numpy(), fw.pos, err_msg=f"positions, step {k}")
    np.testing.assert_array_equal(st.velocities_numpy(), fw.vel, err_msg=f"velocities, step {k}")
    np.testing.assert_array_equal(st.densities_numpy(), fw.rho, err_msg=f"densities, step {k}")
    p, lv, av = world_state(world)
    np.testing.assert_array_equal(p, fw.rigid.pose7, err_msg=f"body poses, step {k}")
    np.testing.assert_array_equal(lv, fw.rigid.lin, err_msg=f"body velocities, step {k}")
    np.testing.assert_array_equal(av, fw.rigid.ang, err_msg=f"body angular velocities, step {k}")


def _small_coupled_world():
    """A 2048-particle dam break (configured parameters) in a 1.2 x 1 x 1.2 m
    box with a ground plane, two spheres and two boxes dropping into the fluid
    and onto each other, one box resting on the floor beside the fluid, and a
    static box standing in the fluid."""
    sc = scenes.dam_break(16, 8, 16, (1.2, 1.0, 1.2), seed=31, steps=50)
    top = float(sc.positions[:, 1].max())
    b = nv.ModelBuilder()
    b.add_ground_plane(y=0.0)

    def at(x, y, z, q=(0.0, 0.0, 0.0, 1.0)):
        return nv.Transform(np.array([x, y, z], F32), np.array(q, F32))

    i = b.add_body(nv.RigidBody.from_sphere(3.0, 0.12), at(0.30, top + 0.06, 0.30))
    b.add_shape(nv.CollisionShape.make_sphere(0.12, i))
    i = b.add_body(nv.RigidBody.from_sphere(1.5, 0.1), at(0.36, top + 0.32, 0.33))
    b.add_shape(nv.CollisionShape.make_sphere(0.1, i))
    he = np.array([0.1, 0.07, 0.12], F32)
    s30 = math.sin(math.pi / 12)
    i = b.add_body(nv.RigidBody.from_box(2.5, he), at(0.60, top + 0.05, 0.55, (0.0, s30, 0.0, math.cos(math.pi / 12))))
    b.add_shape(nv.CollisionShape.make_box(he, i))
    he2 = np.array([0.09, 0.06, 0.08], F32)
    i = b.add_body(nv.RigidBody.from_box(1.2, he2), at(0.62, top + 0.25, 0.58))
    b.add_shape(nv.CollisionShape.make_box(he2, i))
    he3 = np.array([0.08, 0.08, 0.08], F32)
    i = b.add_body(nv.RigidBody.from_box(1.0, he3), at(1.0, 0.085, 1.0))
    b.add_shape(nv.CollisionShape.make_box(he3, i))
    i = b.add_body(nv.RigidBody.make_static(), at(0.15, 0.2, 0.65))
    b.add_shape(nv.CollisionShape.make_box(np.array([0.05, 0.2, 0.05], F32), i))
    world = nv.FluidWorld(b.build(), [scenes.empty_block()], nv.SolverSettings(), scenes.pbf_settings(sc.params),
                          boundary_extent=1.2)
    world.fluid_state.init_arrays(sc.positions, sc.velocities)
    return world, sc


def test_coupled_rollout_boxes_spheres_50_steps_matches_oracle(gpu):
    world, sc = _small_coupled_world()
    _say("case: coupled FluidWorld 50 steps, n=", sc.n, "boundary=", world.num_boundary_particles)
    h = sc.params.kernel_radius
    psi = _initial_psi(world, h)
    fw = _oracle_run(world, sc, psi)
    dt = float(F32(sc.dt))
    contacts = 0
    for k in range(1, 51):
        world.step(dt)
        fw.step(dt, sc.gravity)
        if k == 1:
            np.testing.assert_array_equal(boundary_arrays(world)[2], psi)
        contacts += len(fw.rigid.contacts)
        if k in (1, 2, 5, 10, 20, 30, 40, 50):
            _compare(world, fw, k)
    p, lv, av = world_state(world)
    assert contacts > 0  # bodies touched the floor and each other
    assert np.abs(lv[:5]).max() > 0.1 and np.all(lv[5] == 0)  # the static box never moves


def test_fluid_world_first_step_hands_oracle_forces_to_rigid_step(gpu):
    """One FluidWorld step: besides the fluid state, the forces and torques
    the coupling hands to World::apply_force / apply_torque are the oracle's.
    They are consumed by the rigid step, so the bodies' velocities after it
    pin them: the oracle rigid world fed the oracle's forces must match."""
    world, sc = _small_coupled_world()
    psi = _initial_psi(world, sc.params.kernel_radius)
    fw = _oracle_run(world, sc, psi)
    dt = float(F32(sc.dt))
    world.step(dt)
    fw.step(dt, sc.gravity)
    _compare(world, fw, 1)
    # the same step without the coupling forces differs: the check has teeth
    ow = oracle_rigid(world.model, world.model.initial_transforms, world.solver_settings)
    ow.step(dt)
    assert not np.array_equal(ow.lin, fw.rigid.lin)


def test_config2_500k_fluid_64_bodies_step_matches_oracle(gpu):
    """The coupling workload at its stated size: 500k particles, a sampled
    ground plane over the box and 64 spheres and boxes, one FluidWorld step
    (PBF, boundary density, coupling, host rigid step) against the oracle."""
    world, sc = scenes.coupled_500k_world()
    _say("case: config 2, n=", sc.n, "boundary=", world.num_boundary_particles)
    psi = _initial_psi(world, sc.params.kernel_radius)
    fw = _oracle_run(world, sc, psi)
    dt = float(F32(sc.dt))
    world.step(dt)
    fw.step(dt, sc.gravity)
    np.testing.assert_array_equal(boundary_arrays(world)[2], psi)
    _compare(world, fw, 1)


def test_config3_1024_scene_batch(gpu):
    """The 1024-scene batch (N log-uniform 2k-64k, configured parameters)
    for 3 steps: 16 seeded scenes equal their single-scene contexts and the
    two smallest scenes equal the oracle, bitwise."""
    scs = scenes.batched_1024()
    slots = sum(((s.n + 255) // 256) * 256 for s in scs)
    cells = sum(scenes.window_cells(s) for s in scs)
    _say("case: config 3,", len(scs), "scenes,", sum(s.n for s in scs), "particles")
    b = capi.Batch(slot_capacity=slots, cell_capacity=cells, max_scenes=len(scs))
    ids = [b.add(s.params, s.dt, s.gravity, s.mass, s.positions, s.velocities) for s in scs]
    b.step(3)
    assert not b.status(len(scs)).any()
    rng = np.random.default_rng(33)
    sample = sorted(rng.choice(len(scs), 16, replace=False).tolist())
    for k in sample:
        s = scs[k]
        c = capi.Context(s.params, capacity=s.n)
        c.set_particles(s.positions, s.velocities)
        c.step(s.dt, s.gravity, s.mass, 3)
        want = c.get_particles()
        c.close()
        got = b.get(ids[k])
        for g, w in zip(got, want):
            np.testing.assert_array_equal(g, w, err_msg=f"scene {k}")
    for k in np.argsort([s.n for s in scs])[:2]:
        s = scs[int(k)]
        op, ov, orho, ol = O.pbf_run(s.params, s.positions, s.velocities, s.dt, s.gravity, s.mass, 3)
        gp, gv, gr, gl = b.get(ids[int(k)])
        np.testing.assert_array_equal(gp, op)
        np.testing.assert_array_equal(gv, ov)
        np.testing.assert_array_equal(gr, orho)
        np.testing.assert_array_equal(gl, ol)
    b.close()


def test_batch_index_space_rank_equals_alone(gpu):
    """Six collapsing scenes (configured parameters, 6k-40k particles) in one
    batch with huge_levels 0 (every class spanning two or more blocks ranks in
    index space, each class's row scanned over its own scene's tiles, up to 64
    classes per step and the rest by merge launches) over 40 steps: bitwise
    equal to each scene stepped alone with the default split (the single
    contexts are pinned to the oracle by the single-scene tests)."""
    scs = [scenes.batched_scene(nx, ny, nz, f, seed=700 + k, s_corr="configured")
           for k, (nx, ny, nz, f) in enumerate([(20, 20, 15, 0.3), (40, 25, 40, 0.25), (30, 30, 30, 0.4),
                                                (25, 20, 25, 0.2), (16, 24, 16, 0.3), (34, 20, 30, 0.35)])]
    slots = sum(((s.n + 255) // 256) * 256 for s in scs)
    cells = sum(scenes.window_cells(s) for s in scs)
    _say("case: batch with index-space rank,", len(scs), "scenes,", sum(s.n for s in scs), "particles, 40 steps")
    b = capi.Batch(slot_capacity=slots, cell_capacity=cells, max_scenes=len(scs))
    b.set_rank_split(0)
    ids = [b.add(s.params, s.dt, s.gravity, s.mass, s.positions, s.velocities) for s in scs]
    b.step(40)
    assert not b.status(len(scs)).any()
    for k, s in enumerate(scs):
        c = capi.Context(s.params, capacity=s.n)
        c.set_particles(s.positions, s.velocities)
        c.step(s.dt, s.gravity, s.mass, 40)
        want = c.get_particles()
        c.close()
        got = b.get(ids[k])
        for g, w in zip(got, want):
            np.testing.assert_array_equal(g, w, err_msg=f"scene {k}")
    b.close()


class _Clock:
    t = 0.0

    def __call__(self):
        return self.t


def test_config4_served_full_size_range(gpu):
    """Served rollouts drawn as the served workload draws them (Poisson
    arrivals, seed 4, N log-uniform 4k-256k, 0-8 spheres) with horizons cut
    to 6-10 steps: every request finishes, later arrivals join while others
    run, and a sample (the largest fluid-only request, the smallest and the
    largest coupled ones) is bitwise the same alone as in the mix."""
    reqs = serve.poisson_requests(10, 4.0, seed=4, make_scene=lambda k, rng: scenes.served_request(k, rng),
                                  horizon_range=(6, 10))
    _say("case: config 4,", [(r.scene.n, len(r.scene.spheres)) for r in reqs])
    sizes = [r.scene.n for r in reqs]
    assert min(sizes) >= 4000 and max(sizes) <= 256000
    b = capi.Batch(slot_capacity=sum(((n + 255) // 256) * 256 for n in sizes),
                   cell_capacity=sum(scenes.window_cells(r.scene) for r in reqs), max_scenes=16)
    clock = _Clock()
    srv = serve.Server(serve.FusedBackend(b, 16), hbm_budget_bytes=int(0.8 * 288e9), max_scenes=16,
                       bytes_fn=capi.scene_bytes, clock=clock, health_every=1)
    for r in reqs:
        srv.submit(r)
    for t in range(200):
        clock.t = 0.25 * t
        srv.step()
        if not srv.pending and not srv.active:
            break
    assert all(srv.finished[r.rid].status == serve.DONE for r in reqs)
    assert len({srv.finished[r.rid].admitted_at for r in reqs}) > 3
    lat = srv.latency_percentiles()
    assert lat["samples"] == sum(r.horizon for r in reqs) and lat["p99_ms"] >= lat["p50_ms"] > 0
    fluid = [r for r in reqs if not r.scene.spheres]
    coupled = sorted([r for r in reqs if r.scene.spheres], key=lambda r: r.scene.n)
    sample = ([max(fluid, key=lambda r: r.scene.n)] if fluid else []) + coupled[:1] + coupled[-1:]
    assert len(sample) >= 2
    for r in sample:
        sc, res = r.scene, srv.finished[r.rid].result
        if not sc.spheres:
            c = capi.Context(sc.params, capacity=sc.n)
            c.set_particles(sc.positions, sc.velocities)
            c.step(sc.dt, sc.gravity, sc.mass, r.horizon)
            want = c.get_particles()
            c.close()
            for g, w in zip(res[:4], want):
                np.testing.assert_array_equal(g, w, err_msg=f"request {r.rid}")
            continue
        w = serve.fluid_world_for(sc)
        for _ in range(r.horizon):
            w.step(float(sc.dt))
        st = w.fluid_state
        np.testing.assert_array_equal(res[0], st.positions_numpy(), err_msg=f"request {r.rid}")
        np.testing.assert_array_equal(res[1], st.velocities_numpy(), err_msg=f"request {r.rid}")
        np.testing.assert_array_equal(res[2], st.densities_numpy(), err_msg=f"request {r.rid}")
        for got, want in zip(res[4], [np.array(t.position) for t in w.state.transforms]):
            np.testing.assert_array_equal(got, want)
    b.close()


def _state_hash(pos, vel, rho, lam):
    h = hashlib.sha256()
    for a in (pos, vel, rho, lam):
        h.update(np.ascontiguousarray(a, "<f4").tobytes())
    return h.hexdigest()


@pytest.mark.parametrize("huge_levels,lists", [(6, 0), (2, 1)])
def test_config0_configured_200_steps_against_oracle_rollout(gpu, huge_levels, lists):
    """The 8000-particle dam break with the configured s_corr (k = 0.1,
    dq = 0.2h), the scene the reference's compute_delta_position drives into
    collapse: 200 GPU steps against the oracle's committed rollout
    (tests/golden/make_config0_rollout.py).  Mean density, mean density error
    and kinetic energy every 10 steps within 1 %, the tolerance of a 200-step
    rollout, and the final state's hash equal (the default kernels are
    bitwise).  huge_levels 2: the corner stack (most of the 8000 particles in
    one class, up to 31 blocks) ranks in index space instead of by merge
    launches; with lists 1 the steps build skin lists where the previous
    step's bounds mostly held; the same hash."""
    with open(os.path.join(HERE, "golden", "config0_configured_200.json")) as f:
        gold = json.load(f)
    sc = scenes.dam_break_8k()
    _say("case: config 0 configured 200 steps, huge_levels", huge_levels, "lists", lists)
    c = capi.Context(sc.params, capacity=sc.n)
    c.set_rank_split(huge_levels)
    c.set_list_policy(lists)
    c.set_particles(sc.positions, sc.velocities)
    m = float(F32(sc.mass))
    for row in gold["rows"]:
        c.step(sc.dt, sc.gravity, sc.mass, 10)
        pos, vel, rho, lam = c.get_particles()
        v = vel.astype(np.float64)
        got = {"mean_density": float(rho.astype(np.float64).mean()),
               "mean_density_error": float(np.abs(rho.astype(np.float64) / sc.params.rest_density - 1.0).mean()),
               "kinetic_energy": float(0.5 * m * (v * v).sum())}
        for key, val in got.items():
            assert val == pytest.approx(row[key], rel=0.01), (row["step"], key)
    assert _state_hash(pos, vel, rho, lam) == gold["final_state_sha256"]
    c.close()


def test_query_all_pairs_on_a_dam_break_matches_oracle(gpu):
    """SpatialHashGrid::query_all_pairs (neighbor_search.cpp:57-87) on the
    8000-particle dam break at the kernel radius, and on the reference test's
    500-point set, equals the oracle's pairs in order."""
    sc = scenes.dam_break_8k()
    pts = [p for p in sc.positions]
    grid = nv.SpatialHashGrid(0.1)
    grid.build(pts)
    got = grid.query_all_pairs(pts, 0.1)
    want = O.grid_all_pairs(sc.positions, 0.1, 0.1)
    assert len(got) == want.shape[0] > 8000
    np.testing.assert_array_equal(np.array(got, np.int32), want)
    np.random.seed(42)
    pts = np.random.uniform(0, 1, (500, 3)).astype(F32)
    grid = nv.SpatialHashGrid(0.1)
    grid.build(list(pts))
    got = grid.query_all_pairs(list(pts), 0.1)
    np.testing.assert_array_equal(np.array(got, np.int32).reshape(-1, 2), O.grid_all_pairs(pts, 0.1, 0.1))


@pytest.mark.parametrize("huge_levels", [6, 1])
def test_aux_sorts_with_large_classes_match_oracle(gpu, huge_levels):
    """The auxiliary grids' rank inside large classes (a stack of 3000
    coincident points, 800 points in one cell, classes spanning many 256-slot
    blocks, merged level by level): standalone queries return the oracle's
    bucket lists in order, and coupling over a fluid holding a stack of 2000
    coincident particles next to a body gives the oracle's forces.  With
    huge_levels 1 the context ranks the stack's class (8 blocks) in index
    space in its step and its coupling grid."""
    rng = np.random.default_rng(71)
    pts = np.concatenate([np.full((3000, 3), 0.437, F32), rng.uniform(0.40, 0.49, (800, 3)),
                          rng.uniform(-0.5, 1.5, (4000, 3))]).astype(F32)
    pts = pts[rng.permutation(pts.shape[0])]
    q = np.concatenate([pts[:50], rng.uniform(0.3, 0.6, (50, 3))]).astype(F32)
    counts, lists = capi.grid_query(pts, 0.1, q)
    oc, ol = O.grid_query(pts, 0.1, q)
    np.testing.assert_array_equal(counts, oc)
    np.testing.assert_array_equal(lists, ol)
    # coupling: the committed-fluid grid sorts the fluid by original index
    sc, local, body, poses, static = _stack_scene_with_body(rng)
    h = sc.params.kernel_radius
    c = capi.Context(sc.params, capacity=sc.n)
    c.set_rank_split(huge_levels)
    c.set_particles(sc.positions, sc.velocities)
    c.set_boundary(local, body, None, nbodies=poses.shape[0])
    psi = c.compute_boundary_volumes(poses)
    c.step(sc.dt, sc.gravity, sc.mass, 1)
    c.boundary_world_positions(poses)
    c.apply_boundary_density()
    f, t = c.coupling_forces(poses, static)
    gp, gv, gr, gl = c.get_particles()
    c.close()
    op, ov, orho, _ = O.pbf_run(sc.params, sc.positions, sc.velocities, sc.dt, sc.gravity, sc.mass, 1)
    wpos = O.boundary_world_positions(local, body, poses)
    orho_b = O.boundary_density(op, orho, wpos, psi, h, sc.params.rest_density)
    of, ot = O.coupling(op, orho_b, wpos, psi, body, poses, static, h, sc.params.rest_density, sc.mass)
    np.testing.assert_array_equal(gp, op)
    np.testing.assert_array_equal(gr, orho_b)
    np.testing.assert_array_equal(f, of)
    np.testing.assert_array_equal(t, ot)
    assert np.abs(of).max() > 0


def _stack_scene_with_body(rng):
    """A dam break whose first 2000 particles sit on one point beside a sphere
    (a stack that stays together: equal positions and velocities), and the
    sphere's boundary samples."""
    sc = scenes.dam_break(12, 10, 12, (1.0, 1.0, 1.0), seed=72, steps=1)
    pos = np.concatenate([np.tile(np.array([[0.31, 0.26, 0.33]], F32), (2000, 1)), sc.positions]).astype(F32)
    sc.positions = pos
    sc.velocities = np.zeros_like(pos)
    shapes = [O.make_shape(O.SHAPE_SPHERE, 0, [0.12, 0, 0, 0])]
    local, body = O.sample_boundaries(shapes, 0.05, 1.0)
    poses = np.array([[0.42, 0.30, 0.36, 1, 0, 0, 0]], F32)
    return sc, local, body, poses, np.array([0], np.uint8)
