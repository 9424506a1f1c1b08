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
