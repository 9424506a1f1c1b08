"""Continuous-batching server for fluid rollouts on one GPU.

Requests (a dam-break scene + a horizon in steps) arrive over time.  Between
two device steps the server
  * retires requests that reached their horizon (state read back, slots freed),
  * evicts requests whose scene went bad (NaN / inf, left its cell window)
    without disturbing the others,
  * admits waiting requests in arrival order while the HBM budget, the scene
    slots and the batch capacities allow (a request larger than the whole
    budget is rejected at once),
then advances every admitted request by one step.  Particle and grid state
stay on the device for the whole rollout.  A request's result is bitwise the
same alone or in any mix (per-scene arithmetic does not depend on the batch).

Backends (``add(scene) -> key, step() -> ms, status() -> {key: code},
get(key), remove(key)``):
  * BatchBackend: fluid-only requests in one batch context; every request
    advances in the same launches (npx_batch_step_timed), and a step's
    latency is its device time.
  * FusedBackend: every request in one batch context; the Akinci coupling
    of requests with rigid bodies (boundary density, coupling forces) runs
    for all of them in the same launches (npx_batch_coupling), and their
    rigid bodies step on the host (World), as FluidWorld does.  A round's
    latency is the host time of the synchronised round.
  * MixedBackend: fluid-only requests as above, and requests with rigid
    bodies as separate FluidWorld instances, each on its own device context
    and stream, stepped from a thread pool while the batch step runs.
The scheduling logic takes any object with those methods, which is how the
CPU tests exercise admission, eviction and cancellation without a GPU.
"""
from __future__ import annotations

import collections
import concurrent.futures
import dataclasses
import time
from typing import Callable, Dict, Iterator, List, Optional

import numpy as np

DONE, CANCELLED, REJECTED, FAILED, RUNNING, WAITING, PREEMPTED = (
    "done", "cancelled", "rejected", "failed", "running", "waiting", "preempted")


@dataclasses.dataclass
class Request:
    rid: int
    scene: object                 # scenes.Scene: params, positions, velocities, dt, gravity, mass
    horizon: int                  # steps to run
    arrival: float = 0.0          # arrival time (seconds on the server's clock)
    status: str = WAITING
    steps_done: int = 0
    sid: object = None            # backend key while admitted
    bytes: int = 0
    step_ms: List[float] = dataclasses.field(default_factory=list)
    result: Optional[tuple] = None
    error: str = ""
    admitted_at: float = 0.0
    finished_at: float = 0.0
    snapshot: Optional[dict] = None  # state to resume from (novaphy_amd.snapshot layout)


class Server:
    def __init__(self, backend, hbm_budget_bytes: int, max_scenes: int,
                 bytes_fn: Callable[[object, int], int], clock: Callable[[], float] = time.perf_counter,
                 health_every: int = 10, keep_results: bool = True):
        self.b = backend
        self.budget = int(hbm_budget_bytes)
        self.max_scenes = int(max_scenes)
        self.bytes_fn = bytes_fn
        self.clock = clock
        self.health_every = max(1, int(health_every))
        self.keep_results = bool(keep_results)  # False: drop each result (and scene) once read back
        self.pending: "collections.deque[Request]" = collections.deque()
        self.active: Dict[int, Request] = {}
        self.finished: Dict[int, Request] = {}
        self.preempted: Dict[int, Request] = {}
        self.used = 0
        self.steps = 0
        self.step_ms: List[float] = []
        self.particle_steps = 0
        self.busy_s = 0.0
        self.admit_s = 0.0       # host time spent admitting requests

    # ---------------------------------------------------------- intake
    def submit(self, req: Request) -> None:
        req.bytes = self.bytes_fn(req.scene.params, req.scene.n)
        if req.bytes > self.budget:
            req.status = REJECTED
            req.error = f"needs {req.bytes} B, budget is {self.budget} B"
            self.finished[req.rid] = req
            return
        req.status = WAITING
        self.pending.append(req)

    def cancel(self, rid: int) -> bool:
        for r in list(self.pending):
            if r.rid == rid:
                self.pending.remove(r)
                r.status = CANCELLED
                self.finished[rid] = r
                return True
        r = self.active.get(rid)
        if r is not None:
            self._retire(r, CANCELLED, read=False)
            return True
        return False

    def preempt(self, rid: int) -> bool:
        """Takes a running request off the device: its state is kept as a
        snapshot (novaphy_amd.snapshot layout; save it with
        snapshot.save to survive a restart) and its slots and budget are
        freed.  resume() queues it again; it continues bit for bit."""
        r = self.active.get(rid)
        if r is None:
            return False
        r.snapshot = self.b.snapshot(r.sid)
        self.b.remove(r.sid)
        self.used -= r.bytes
        del self.active[rid]
        r.status = PREEMPTED
        self.preempted[rid] = r
        return True

    def resume(self, rid: int, snapshot: Optional[dict] = None, front: bool = True) -> bool:
        """Queues a preempted request again (ahead of the others with
        front=True), optionally from a snapshot loaded from disk."""
        r = self.preempted.pop(rid, None)
        if r is None:
            return False
        if snapshot is not None:
            r.snapshot = snapshot
        r.status = WAITING
        r.arrival = min(r.arrival, self.clock())
        if front:
            self.pending.appendleft(r)
        else:
            self.pending.append(r)
        return True

    # ---------------------------------------------------------- scheduling
    def _retire(self, r: Request, status: str, read: bool = True) -> None:
        if read:
            r.result = self.b.get(r.sid)
        self.b.remove(r.sid)
        self.used -= r.bytes
        r.status = status
        r.finished_at = self.clock()
        del self.active[r.rid]
        self.finished[r.rid] = r
        if not self.keep_results:
            r.result = None
            r.scene = None

    def _admit(self, now: float) -> None:
        while self.pending and self.pending[0].arrival <= now:
            r = self.pending[0]
            if len(self.active) >= self.max_scenes or self.used + r.bytes > self.budget:
                return                       # FIFO: later arrivals wait behind the head
            try:
                sid = self.b.restore(r.scene, r.snapshot) if r.snapshot is not None else self.b.add(r.scene)
            except Exception as e:            # capacity exhausted on the device side
                if "capacity" in str(e) or "slots" in str(e):
                    return
                self.pending.popleft()
                r.status = FAILED
                r.error = str(e)
                self.finished[r.rid] = r
                continue
            self.pending.popleft()
            r.sid = sid
            r.snapshot = None
            r.status = RUNNING
            r.admitted_at = now
            self.used += r.bytes
            self.active[r.rid] = r

    def _check_health(self) -> None:
        st = self.b.status()
        for r in list(self.active.values()):
            if st.get(r.sid, 0):
                r.error = f"scene status {int(st[r.sid])} (1 = left its cell window, 2 = non-finite state)"
                self._retire(r, FAILED, read=False)

    def step(self) -> bool:
        """One scheduling round + one device step.  Returns False when idle
        (nothing active and nothing that has arrived)."""
        now = self.clock()
        ta = time.perf_counter()
        self._admit(now)
        self.admit_s += time.perf_counter() - ta
        if not self.active:
            return False
        t0 = time.perf_counter()
        ms = float(self.b.step())
        self.busy_s += time.perf_counter() - t0
        self.steps += 1
        self.step_ms.append(ms)
        for r in list(self.active.values()):
            r.steps_done += 1
            r.step_ms.append(ms)
            self.particle_steps += r.scene.n
        if self.steps % self.health_every == 0:
            self._check_health()
        for r in list(self.active.values()):
            if r.steps_done >= r.horizon:
                self._retire(r, DONE)
        return True

    def run(self, until_empty: bool = True, max_rounds: int = 10 ** 9) -> None:
        rounds = 0
        while rounds < max_rounds:
            rounds += 1
            if not self.step():
                if not self.pending:
                    if until_empty:
                        return
                else:
                    # idle until the next arrival on the server's clock
                    nxt = self.pending[0].arrival
                    wait = nxt - self.clock()
                    if wait > 0:
                        time.sleep(min(wait, 0.05))

    # ---------------------------------------------------------- report
    def latency_percentiles(self):
        samples = np.concatenate([np.asarray(r.step_ms, np.float64) for r in self.finished.values()
                                  if r.step_ms] or [np.zeros(0)])
        if samples.size == 0:
            return {"p50_ms": None, "p99_ms": None, "samples": 0}
        return {"p50_ms": float(np.percentile(samples, 50)), "p99_ms": float(np.percentile(samples, 99)),
                "samples": int(samples.size)}


class BatchBackend:
    """Fluid-only requests in one batch context (capi.Batch)."""

    def __init__(self, batch, max_scenes: int):
        self.b = batch
        self.max_scenes = int(max_scenes)

    def add(self, sc):
        return self.b.add(sc.params, sc.dt, sc.gravity, sc.mass, sc.positions, sc.velocities)

    def remove(self, key):
        self.b.remove(key)

    def get(self, key):
        return self.b.get(key)

    def snapshot(self, key) -> dict:
        p, v, r, l = self.b.get(key)
        return {"positions": p, "velocities": v, "densities": r, "lambdas": l}

    def restore(self, sc, snap):
        return self.b.add(sc.params, sc.dt, sc.gravity, sc.mass, snap["positions"], snap["velocities"])

    def step(self) -> float:
        return float(self.b.step_timed(1)[0])

    def status(self) -> Dict[object, int]:
        st = self.b.status(self.max_scenes)
        return {k: int(v) for k, v in enumerate(st) if v}


def fluid_world_for(sc, plane_extent: Optional[float] = None):
    """FluidWorld of a served request: the scene's fluid particles, its rigid
    spheres and a ground plane at y = 0 for them to land on."""
    import novaphy_amd as nv
    b = nv.ModelBuilder()
    b.add_ground_plane(y=0.0)
    for sp in getattr(sc, "spheres", []):
        mass = float(sp.density * 4.0 / 3.0 * np.pi * sp.radius ** 3)
        t = nv.Transform.from_translation(np.array(sp.position, np.float32))
        i = b.add_body(nv.RigidBody.from_sphere(mass, float(sp.radius)), t)
        b.add_shape(nv.CollisionShape.make_sphere(float(sp.radius), i))
    model = b.build()
    prm = nv.PBFSettings()
    p = sc.params
    for f in ("rest_density", "kernel_radius", "particle_radius", "solver_iterations", "epsilon",
              "xsph_viscosity", "vorticity_epsilon", "corr_delta_q", "corr_k", "corr_n", "use_domain_bounds"):
        setattr(prm, f, getattr(p, f))
    prm.domain_lower = np.array(p.domain_lower, np.float32)
    prm.domain_upper = np.array(p.domain_upper, np.float32)
    blk = nv.FluidBlockDef()
    blk.particle_spacing = 2.0 * p.particle_radius
    blk.lower = np.zeros(3, np.float32)
    blk.upper = np.zeros(3, np.float32)
    if plane_extent is None:
        plane_extent = float(max(p.domain_upper[0], p.domain_upper[2]))
    world = nv.FluidWorld(model, [blk], nv.SolverSettings(), prm, boundary_extent=plane_extent)
    world.fluid_state.init_arrays(sc.positions, sc.velocities)
    return world


def _model_for(sc):
    """Rigid model of a served request: its spheres and a ground plane at
    y = 0 (the same model fluid_world_for builds)."""
    import novaphy_amd as nv
    b = nv.ModelBuilder()
    b.add_ground_plane(y=0.0)
    for sp in getattr(sc, "spheres", []):
        mass = float(sp.density * 4.0 / 3.0 * np.pi * sp.radius ** 3)
        t = nv.Transform.from_translation(np.array(sp.position, np.float32))
        i = b.add_body(nv.RigidBody.from_sphere(mass, float(sp.radius)), t)
        b.add_shape(nv.CollisionShape.make_sphere(float(sp.radius), i))
    return b.build()


def _poses7(transforms):
    out = np.zeros((len(transforms), 7), np.float32)
    for k, t in enumerate(transforms):
        out[k, :3] = np.array(t.position, np.float32)
        x, y, z, w = np.array(t.rotation, np.float32)  # xyzw
        out[k, 3:] = (w, x, y, z)
    return out


class FusedBackend:
    """Every request in one batch context.  A request with rigid spheres
    gets its Akinci boundary in the batch (npx_scene_set_boundary, volumes
    computed as FluidWorld computes them) and a host World for its bodies;
    each round is one batch step, one npx_batch_coupling over all coupled
    scenes and the host rigid steps (FluidWorld::step order).

    device_sampling: sample the boundary particles on the device
    (npx_sample_boundaries) instead of with the host samplers FluidWorld
    uses; the samples are the same bit for bit, and the arrays skip the
    host's per-particle objects."""

    def __init__(self, batch, max_scenes: int, device_sampling: bool = False):
        self.batch = BatchBackend(batch, max_scenes)
        self.coupled: Dict[int, tuple] = {}   # sid -> (World, nbodies, is_static, dt)
        self.device_sampling = device_sampling
        # host seconds per part of a round: the batch step (synchronised), the
        # gathering of the body poses, the coupling call (with its readback),
        # the host rigid steps
        self.phase_s = {"fluid": 0.0, "poses": 0.0, "coupling": 0.0, "rigid": 0.0}

    def _attach(self, sid, sc, snap=None):
        import novaphy_amd as nv
        from novaphy_amd import capi
        model = _model_for(sc)
        p = sc.params
        extent = float(max(p.domain_upper[0], p.domain_upper[2]))
        if self.device_sampling:
            shapes, nsh = capi.shapes_from_model(model)
            local, body = capi.sample_boundaries(shapes, nsh, 2.0 * p.particle_radius, extent)
        else:
            bnd = nv.sample_model_boundaries(model, 2.0 * p.particle_radius, extent)
            local = np.array([np.array(b.local_position, np.float32) for b in bnd], np.float32).reshape(-1, 3)
            body = np.array([b.body_index for b in bnd], np.int32)
        nbod = model.num_bodies
        # volumes at the model's initial transforms, on a scratch context as
        # FluidWorld computes them (boundary.cpp:152-176)
        tmp = capi.Context(p, capacity=1)
        tmp.set_boundary(local, body, None, nbod)
        psi = tmp.compute_boundary_volumes(_poses7(model.initial_transforms))
        tmp.close()
        static = np.array([b.is_static() for b in model.bodies], np.uint8)
        self.batch.b.set_boundary(sid, local, body, psi, nbod, static)
        world = nv.World(model, nv.SolverSettings())
        if snap is not None and "body_positions" in snap:
            ss = world.state
            for i in range(len(snap["body_positions"])):
                ss.set_transform(i, nv.Transform(snap["body_positions"][i], snap["body_rotations"][i]))
                ss.set_linear_velocity(i, snap["body_linear_velocities"][i])
                ss.set_angular_velocity(i, snap["body_angular_velocities"][i])
        self.coupled[sid] = (world, nbod, static, float(sc.dt))

    def add(self, sc):
        sid = self.batch.add(sc)
        if getattr(sc, "spheres", None):
            self._attach(sid, sc)
        return sid

    def restore(self, sc, snap):
        sid = self.batch.restore(sc, snap)
        if getattr(sc, "spheres", None):
            self._attach(sid, sc, snap)
        return sid

    def remove(self, key):
        self.coupled.pop(key, None)
        self.batch.remove(key)

    def get(self, key):
        res = self.batch.get(key)
        if key in self.coupled:
            w = self.coupled[key][0]
            res = tuple(res) + ([np.array(t.position) for t in w.state.transforms],)
        return res

    def snapshot(self, key) -> dict:
        snap = self.batch.snapshot(key)
        if key in self.coupled:
            ss = self.coupled[key][0].state
            snap.update({
                "body_positions": np.array([np.array(t.position) for t in ss.transforms], np.float32),
                "body_rotations": np.array([np.array(t.rotation) for t in ss.transforms], np.float32),
                "body_linear_velocities": np.array([np.array(v) for v in ss.linear_velocities], np.float32),
                "body_angular_velocities": np.array([np.array(v) for v in ss.angular_velocities], np.float32)})
        return snap

    def step(self) -> float:
        t0 = time.perf_counter()
        self.batch.step()
        t1 = time.perf_counter()
        self.phase_s["fluid"] += t1 - t0
        if self.coupled:
            import novaphy_amd as nv
            keys = sorted(self.coupled)  # the body table order of npx_batch_coupling
            worlds = [self.coupled[k][0] for k in keys]
            poses = nv.gather_poses7(worlds)
            tp = time.perf_counter()
            self.phase_s["poses"] += tp - t1
            out = self.batch.b.coupling(poses)
            t2 = time.perf_counter()
            self.phase_s["coupling"] += t2 - tp
            # FluidWorld::step's rules in one host call: static bodies and
            # bodies whose force and torque both have a zero float32 squared
            # norm get nothing, then each world takes its rigid step
            nv.apply_forces_and_step(worlds, out, [self.coupled[k][3] for k in keys])
            self.phase_s["rigid"] += time.perf_counter() - t2
        return (time.perf_counter() - t0) * 1e3

    def status(self) -> Dict[object, int]:
        return self.batch.status()


class MixedBackend:
    """Fluid-only requests in a batch context; requests with rigid bodies as
    FluidWorld instances (own context and stream each) stepped concurrently
    from a thread pool while the batch step runs."""

    def __init__(self, batch, max_scenes: int, threads: int = 8):
        self.batch = BatchBackend(batch, max_scenes)
        self.n_batch = 0
        self.worlds: Dict[int, tuple] = {}
        self.next_world = 0
        self.failed: Dict[object, int] = {}
        self.pool = concurrent.futures.ThreadPoolExecutor(max_workers=max(1, threads))

    def add(self, sc):
        if not getattr(sc, "spheres", None):
            key = ("b", self.batch.add(sc))
            self.n_batch += 1
            return key
        key = ("w", self.next_world)
        self.next_world += 1
        self.worlds[key[1]] = (fluid_world_for(sc), float(sc.dt))
        return key

    def remove(self, key):
        self.failed.pop(key, None)
        if key[0] == "b":
            self.batch.remove(key[1])
            self.n_batch -= 1
        else:
            self.worlds.pop(key[1], None)

    def get(self, key):
        if key[0] == "b":
            return self.batch.get(key[1])
        w, _ = self.worlds[key[1]]
        st = w.fluid_state
        return (st.positions_numpy(), st.velocities_numpy(), st.densities_numpy(), st.lambdas_numpy(),
                [np.array(t.position) for t in w.state.transforms])

    def snapshot(self, key) -> dict:
        if key[0] == "b":
            return self.batch.snapshot(key[1])
        w, _ = self.worlds[key[1]]
        st = w.fluid_state
        ss = w.state
        return {"positions": st.positions_numpy(), "velocities": st.velocities_numpy(),
                "densities": st.densities_numpy(), "lambdas": st.lambdas_numpy(),
                "body_positions": np.array([np.array(t.position) for t in ss.transforms], np.float32),
                "body_rotations": np.array([np.array(t.rotation) for t in ss.transforms], np.float32),
                "body_linear_velocities": np.array([np.array(v) for v in ss.linear_velocities], np.float32),
                "body_angular_velocities": np.array([np.array(v) for v in ss.angular_velocities], np.float32)}

    def restore(self, sc, snap):
        if not getattr(sc, "spheres", None):
            key = ("b", self.batch.restore(sc, snap))
            self.n_batch += 1
            return key
        import novaphy_amd as nv
        w = fluid_world_for(sc)  # boundary particles and psi as at the start of the rollout
        w.fluid_state.init_arrays(snap["positions"], snap["velocities"])
        ss = w.state
        for i in range(len(snap.get("body_positions", []))):
            ss.set_transform(i, nv.Transform(snap["body_positions"][i], snap["body_rotations"][i]))
            ss.set_linear_velocity(i, snap["body_linear_velocities"][i])
            ss.set_angular_velocity(i, snap["body_angular_velocities"][i])
        key = ("w", self.next_world)
        self.next_world += 1
        self.worlds[key[1]] = (w, float(sc.dt))
        return key

    def _step_world(self, k):
        w, dt = self.worlds[k]
        try:
            w.step(dt)
        except Exception:
            self.failed[("w", k)] = 2

    def step(self) -> float:
        t0 = time.perf_counter()
        futs = [self.pool.submit(self._step_world, k) for k in list(self.worlds)]
        if self.n_batch:
            self.batch.step()
        for f in futs:
            f.result()
        return (time.perf_counter() - t0) * 1e3

    def status(self) -> Dict[object, int]:
        out = {("b", k): v for k, v in self.batch.status().items()}
        out.update(self.failed)
        return out

    def close(self):
        self.pool.shutdown(wait=True)
        self.worlds.clear()


def poisson_stream(n: int, rate_hz: float, seed: int, make_scene: Callable[[int, np.random.Generator], object],
                   horizon_range=(50, 500)) -> Iterator[Request]:
    """n requests with exponential inter-arrival times at rate_hz (seeded),
    each scene built when its request is drawn."""
    rng = np.random.default_rng(seed)
    t = 0.0
    for k in range(n):
        t += rng.exponential(1.0 / rate_hz)
        h = int(rng.integers(horizon_range[0], horizon_range[1] + 1))
        yield Request(rid=k, scene=make_scene(k, rng), horizon=h, arrival=t)


def poisson_requests(n: int, rate_hz: float, seed: int, make_scene: Callable[[int, np.random.Generator], object],
                     horizon_range=(50, 500)) -> List[Request]:
    """The whole stream of poisson_stream as a list."""
    return list(poisson_stream(n, rate_hz, seed, make_scene, horizon_range))
