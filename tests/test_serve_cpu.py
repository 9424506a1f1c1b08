"""Scheduling logic of the continuous-batching server on the CPU, with a test
double standing in for the device batch (admission by HBM budget, FIFO order,
rejection, cancellation, eviction of a failing scene, horizons, latency
percentiles).  The device path itself is covered by the GPU tests."""
import types

import numpy as np

from novaphy_amd import serve


class FakeBatch:
    def __init__(self, capacity_scenes=100):
        self.scenes = {}
        self.next = 0
        self.bad = set()
        self.cap = capacity_scenes
        self.added = []

    def add(self, sc):
        if len(self.scenes) >= self.cap:
            raise RuntimeError("npx error -3: all scene slots are in use")
        sid = self.next
        self.next += 1
        self.scenes[sid] = np.asarray(sc.positions).shape[0]
        self.added.append(sid)
        return sid

    def remove(self, sid):
        del self.scenes[sid]

    def get(self, sid):
        return ("state", sid)

    def snapshot(self, sid):
        return {"positions": np.full((self.scenes[sid], 3), float(sid), np.float32)}

    def restore(self, sc, snap):
        self.restored = getattr(self, "restored", []) + [snap]
        return self.add(sc)

    def step(self):
        return 2.0 + 0.01 * len(self.scenes)

    def status(self):
        return {s: 2 for s in self.bad if s in self.scenes}


def _scene(n):
    return types.SimpleNamespace(params=None, positions=np.zeros((n, 3), np.float32), velocities=None,
                                 dt=0.01, gravity=(0, -9.81, 0), mass=0.1, n=n)


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t


def test_admission_budget_fifo_and_horizons():
    b = FakeBatch()
    clk = Clock()
    srv = serve.Server(b, hbm_budget_bytes=1000, max_scenes=8, bytes_fn=lambda p, n: n, clock=clk)
    sizes = [400, 500, 300, 200]
    for k, n in enumerate(sizes):
        srv.submit(serve.Request(rid=k, scene=_scene(n), horizon=3 + k, arrival=0.0))
    srv.step()
    # 400 + 500 fit; 300 would exceed 1000 -> waits (FIFO keeps 200 behind it)
    assert sorted(srv.active) == [0, 1] and [r.rid for r in srv.pending] == [2, 3]
    srv.run()
    assert all(srv.finished[k].status == serve.DONE for k in range(4))
    assert [srv.finished[k].steps_done for k in range(4)] == [3, 4, 5, 6]
    assert srv.used == 0 and not srv.active
    lat = srv.latency_percentiles()
    assert lat["samples"] == 3 + 4 + 5 + 6 and lat["p50_ms"] >= 2.0


def test_rejection_cancellation_and_eviction():
    b = FakeBatch()
    clk = Clock()
    srv = serve.Server(b, hbm_budget_bytes=1000, max_scenes=8, bytes_fn=lambda p, n: n, clock=clk,
                       health_every=1)
    srv.submit(serve.Request(rid=0, scene=_scene(5000), horizon=5))
    assert srv.finished[0].status == serve.REJECTED
    srv.submit(serve.Request(rid=1, scene=_scene(100), horizon=50))
    srv.submit(serve.Request(rid=2, scene=_scene(100), horizon=50))
    srv.submit(serve.Request(rid=3, scene=_scene(100), horizon=50, arrival=5.0))
    srv.step()
    assert srv.cancel(3) and srv.finished[3].status == serve.CANCELLED
    b.bad.add(srv.active[2].sid)   # request 2's scene turns non-finite
    srv.step()
    assert srv.finished[2].status == serve.FAILED and "non-finite" in srv.finished[2].error
    assert 1 in srv.active
    assert srv.cancel(1) and srv.finished[1].status == serve.CANCELLED
    assert srv.used == 0 and not b.scenes


def test_arrivals_wait_for_their_time():
    b = FakeBatch()
    clk = Clock()
    srv = serve.Server(b, hbm_budget_bytes=10 ** 9, max_scenes=4, bytes_fn=lambda p, n: n, clock=clk)
    srv.submit(serve.Request(rid=0, scene=_scene(10), horizon=2, arrival=0.0))
    srv.submit(serve.Request(rid=1, scene=_scene(10), horizon=2, arrival=1.0))
    srv.step()
    assert list(srv.active) == [0]
    clk.t = 1.5
    srv.step()
    assert 1 in srv.active


def test_poisson_requests_seeded():
    mk = lambda k, rng: _scene(int(rng.integers(10, 20)))
    a = serve.poisson_requests(50, 100.0, seed=4, make_scene=mk)
    c = serve.poisson_requests(50, 100.0, seed=4, make_scene=mk)
    assert [r.arrival for r in a] == [r.arrival for r in c]
    assert all(50 <= r.horizon <= 500 for r in a)
    assert all(b.arrival > a_.arrival for a_, b in zip(a, a[1:]))


def test_poisson_stream_is_the_list_lazily():
    """The lazy stream draws the same requests (arrivals, horizons, scene
    sizes) as the list, and builds each scene only when it is drawn."""
    made = []

    def mk(k, rng):
        made.append(k)
        return _scene(int(rng.integers(10, 20)))
    a = serve.poisson_requests(30, 10.0, seed=7, make_scene=mk)
    made.clear()
    it = serve.poisson_stream(30, 10.0, seed=7, make_scene=mk)
    first = next(it)
    assert made == [0]
    b = [first] + list(it)
    assert [(r.rid, r.arrival, r.horizon, r.scene.n) for r in a] == [(r.rid, r.arrival, r.horizon, r.scene.n)
                                                                     for r in b]


def test_server_drops_results_when_asked():
    """keep_results=False: a finished request keeps its status, step count and
    latencies but not its result or scene."""
    clk = Clock()
    srv = serve.Server(FakeBatch(), hbm_budget_bytes=1000, max_scenes=8, bytes_fn=lambda p, n: n, clock=clk,
                       keep_results=False)
    srv.submit(serve.Request(rid=0, scene=_scene(10), horizon=2))
    srv.step()
    srv.step()
    r = srv.finished[0]
    assert r.status == serve.DONE and r.steps_done == 2 and len(r.step_ms) == 2
    assert r.result is None and r.scene is None
    assert srv.latency_percentiles()["samples"] == 2


def test_preempt_and_resume_keep_progress():
    b = FakeBatch()
    clk = Clock()
    srv = serve.Server(b, hbm_budget_bytes=1000, max_scenes=8, bytes_fn=lambda p, n: n, clock=clk)
    srv.submit(serve.Request(rid=0, scene=_scene(100), horizon=6))
    srv.submit(serve.Request(rid=1, scene=_scene(100), horizon=3))
    srv.step()
    srv.step()
    assert srv.preempt(0) and srv.finished.get(0) is None and srv.used == 100
    assert srv.active[1].steps_done == 2 and srv.preempted[0].steps_done == 2
    srv.step()                                   # request 1 finishes alone
    assert srv.finished[1].status == serve.DONE
    assert srv.resume(0)
    srv.run()
    r = srv.finished[0]
    assert r.status == serve.DONE and r.steps_done == 6 and len(r.step_ms) == 6
    assert len(b.restored) == 1 and b.restored[0]["positions"].shape == (100, 3)
    assert srv.used == 0


def test_snapshot_file_round_trip(tmp_path):
    from novaphy_amd import snapshot
    rng = np.random.default_rng(0)
    snap = {"positions": rng.random((7, 3), np.float32), "velocities": rng.random((7, 3), np.float32),
            "densities": rng.random(7, np.float32), "lambdas": rng.random(7, np.float32),
            "body_positions": rng.random((2, 3), np.float32), "body_rotations": rng.random((2, 4), np.float32),
            "body_linear_velocities": rng.random((2, 3), np.float32),
            "body_angular_velocities": rng.random((2, 3), np.float32)}
    path = tmp_path / "s.npz"
    snapshot.save(path, snap, steps_done=42, meta={"rid": 3})
    got, steps, meta = snapshot.load(path)
    assert steps == 42 and meta == {"rid": 3} and set(got) == set(snap)
    for k in snap:
        np.testing.assert_array_equal(got[k], snap[k])
    np.savez(tmp_path / "bad.npz", format=np.array("other"))
    import pytest
    with pytest.raises(ValueError):
        snapshot.load(tmp_path / "bad.npz")
