"""Served rollouts on the GPU: requests that join and leave a running batch
(continuous batching) end bitwise equal to the same request run alone, for
fluid-only requests (one batch context) and for requests with rigid spheres
(Akinci coupling and the host rigid pipeline), both with the coupling fused
into the batch (FusedBackend) and with separate FluidWorld contexts stepped
concurrently (MixedBackend)."""
import numpy as np
import pytest

from novaphy_amd import capi, scenes, serve

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t


def _requests(n_fluid, n_coupled, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(n_fluid + n_coupled):
        sc = scenes.served_request(k, rng, lo=2000, hi=6000, max_spheres=0 if k < n_fluid else 2)
        if k >= n_fluid and not sc.spheres:
            sc.spheres = [scenes.Sphere(radius=0.12, density=800.0,
                                        position=(0.2, float(sc.positions[:, 1].max()) + 0.4, 0.2))]
        reqs.append(serve.Request(rid=k, scene=sc, horizon=int(rng.integers(4, 10)), arrival=float(k % 3)))
    return reqs


def _batch_for(reqs):
    slots = sum(((r.scene.n + 255) // 256) * 256 for r in reqs)
    cells = sum(capi.scene_bytes(r.scene.params, 1) for r in reqs)
    return capi.Batch(slot_capacity=slots, cell_capacity=cells, max_scenes=8)


def _run(srv, clock, rounds=40):
    for t in range(rounds):
        clock.t = float(t)
        srv.step()


def test_served_fluid_requests_equal_alone(gpu):
    reqs = _requests(5, 0, seed=21)
    _say("case: served fluid-only requests", [r.scene.n for r in reqs], "horizons", [r.horizon for r in reqs])
    clock = Clock()
    b = _batch_for(reqs)
    srv = serve.Server(serve.BatchBackend(b, 8), hbm_budget_bytes=1 << 40, max_scenes=8,
                       bytes_fn=capi.scene_bytes, clock=clock, health_every=1)
    for r in reqs:
        srv.submit(r)
    _run(srv, clock)
    assert all(srv.finished[r.rid].status == serve.DONE for r in reqs)
    # requests joined at rounds 0, 1 and 2 while others were running
    assert len({srv.finished[r.rid].admitted_at for r in reqs}) == 3
    for r in reqs:
        sc = r.scene
        c = capi.Context(sc.params, capacity=sc.n)
        c.set_particles(sc.positions, sc.velocities)
        c.step(sc.dt, sc.gravity, sc.mass, r.horizon)
        ap, av, ar, al = c.get_particles()
        c.close()
        bp, bv, br, bl = srv.finished[r.rid].result
        np.testing.assert_array_equal(bp, ap)
        np.testing.assert_array_equal(bv, av)
        np.testing.assert_array_equal(br, ar)
    lat = srv.latency_percentiles()
    assert lat["samples"] == sum(r.horizon for r in reqs)
    b.close()


def _backend(kind, b):
    return serve.FusedBackend(b, 8) if kind == "fused" else serve.MixedBackend(b, 8, threads=4)


@pytest.mark.parametrize("kind", ["fused", "mixed"])
def test_served_coupled_requests_equal_alone(gpu, kind):
    reqs = _requests(2, 3, seed=22)
    _say("case: served coupled requests,", kind, [(r.scene.n, len(r.scene.spheres)) for r in reqs])
    clock = Clock()
    b = _batch_for(reqs)
    be = _backend(kind, b)
    srv = serve.Server(be, hbm_budget_bytes=1 << 40, max_scenes=8, bytes_fn=capi.scene_bytes, clock=clock,
                       health_every=1)
    for r in reqs:
        srv.submit(r)
    _run(srv, clock)
    if kind == "mixed":
        be.close()
    assert all(srv.finished[r.rid].status == serve.DONE for r in reqs), \
        [(r.rid, srv.finished[r.rid].status, srv.finished[r.rid].error) for r in reqs]
    for r in reqs:
        sc = r.scene
        res = srv.finished[r.rid].result
        if not sc.spheres:
            c = capi.Context(sc.params, capacity=sc.n)
            c.set_particles(sc.positions, sc.velocities)
            c.step(sc.dt, sc.gravity, sc.mass, r.horizon)
            ap, av, ar, al = c.get_particles()
            c.close()
            np.testing.assert_array_equal(res[0], ap)
            np.testing.assert_array_equal(res[1], av)
            continue
        w = serve.fluid_world_for(sc)
        for _ in range(r.horizon):
            w.step(float(sc.dt))
        st = w.fluid_state
        np.testing.assert_array_equal(res[0], st.positions_numpy())
        np.testing.assert_array_equal(res[1], st.velocities_numpy())
        np.testing.assert_array_equal(res[2], st.densities_numpy())
        for got, want in zip(res[4], [np.array(t.position) for t in w.state.transforms]):
            np.testing.assert_array_equal(got, want)
    b.close()


def test_device_sampled_requests_equal_alone(gpu):
    """Fused serving with the boundary particles sampled on the device
    (npx_sample_boundaries): each coupled request run in the mix ends bitwise
    equal to the same request served alone, and to the same request served
    with the host samplers FluidWorld uses."""
    reqs = _requests(1, 3, seed=24)
    _say("case: device-sampled coupled requests", [(r.scene.n, len(r.scene.spheres)) for r in reqs])

    def serve_all(rs, device=True):
        clock = Clock()
        b = _batch_for(rs)
        srv = serve.Server(serve.FusedBackend(b, 8, device_sampling=device), hbm_budget_bytes=1 << 40, max_scenes=8,
                           bytes_fn=capi.scene_bytes, clock=clock, health_every=1)
        for r in rs:
            srv.submit(r)
        _run(srv, clock)
        out = {r.rid: srv.finished[r.rid] for r in rs}
        b.close()
        return out

    mix = serve_all(reqs)
    host = serve_all([serve.Request(rid=r.rid, scene=r.scene, horizon=r.horizon, arrival=r.arrival) for r in reqs],
                     device=False)
    for r in reqs:
        alone = serve_all([serve.Request(rid=r.rid, scene=r.scene, horizon=r.horizon, arrival=0.0)])[r.rid]
        for got in (mix[r.rid], host[r.rid]):
            assert got.status == serve.DONE and alone.status == serve.DONE
            for a, g in zip(alone.result[:3], got.result[:3]):
                np.testing.assert_array_equal(g, a)
            if r.scene.spheres:
                for a, g in zip(alone.result[4], got.result[4]):
                    np.testing.assert_array_equal(g, a)


@pytest.mark.parametrize("kind", ["fused", "mixed"])
def test_preempted_requests_continue_bitwise(gpu, tmp_path, kind):
    """Requests taken off the device mid-rollout (snapshot written to disk and
    read back) and resumed end bitwise equal to an uninterrupted run, for a
    fluid-only request and one with rigid spheres."""
    from novaphy_amd import snapshot
    reqs = _requests(1, 1, seed=23)
    for r in reqs:
        r.arrival = 0.0
        r.horizon = 8
    _say("case: preempt / resume,", kind, [(r.scene.n, len(r.scene.spheres)) for r in reqs])
    clock = Clock()
    b = _batch_for(reqs)
    be = _backend(kind, b)
    srv = serve.Server(be, hbm_budget_bytes=1 << 40, max_scenes=8, bytes_fn=capi.scene_bytes, clock=clock)
    for r in reqs:
        srv.submit(r)
    _run(srv, clock, rounds=3)
    for r in reqs:
        assert srv.preempt(r.rid)
        path = tmp_path / f"req{r.rid}.npz"
        snapshot.save(path, srv.preempted[r.rid].snapshot, steps_done=srv.preempted[r.rid].steps_done)
        snap, done, _ = snapshot.load(path)
        assert done == 3
        assert srv.resume(r.rid, snapshot=snap)
    _run(srv, clock, rounds=20)
    if kind == "mixed":
        be.close()
    for r in reqs:
        got = srv.finished[r.rid]
        assert got.status == serve.DONE and got.steps_done == 8
        sc = r.scene
        if not sc.spheres:
            c = capi.Context(sc.params, capacity=sc.n)
            c.set_particles(sc.positions, sc.velocities)
            c.step(sc.dt, sc.gravity, sc.mass, 8)
            ap, av, _, _ = c.get_particles()
            c.close()
            np.testing.assert_array_equal(got.result[0], ap)
            np.testing.assert_array_equal(got.result[1], av)
        else:
            w = serve.fluid_world_for(sc)
            for _ in range(8):
                w.step(float(sc.dt))
            np.testing.assert_array_equal(got.result[0], w.fluid_state.positions_numpy())
            np.testing.assert_array_equal(got.result[1], w.fluid_state.velocities_numpy())
            for g, t in zip(got.result[4], w.state.transforms):
                np.testing.assert_array_equal(g, np.array(t.position))
    b.close()
