"""ctypes binding of libnpx_hip.so (the C ABI declared in include/novaphy_hip.h).

This is the thinnest host layer over the HIP path: numpy arrays in, numpy
arrays out, errors raised as exceptions with the library's message.  It fails
loudly when the library is missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnpx_hip.so")

NPX_OK = 0
NPX_EINVAL = -1
NPX_EDEVICE = -2
NPX_ENOMEM = -3
NPX_EDOMAIN = -4
NPX_ESTATE = -5
NPX_EINTERNAL = -6

FLAG_FAST_MATH = 1
FLAG_NO_GRAPH = 2

STATUS_WINDOW = 1
STATUS_NONFINITE = 2

# Every symbol of include/novaphy_hip.h.
SYMBOLS = (
    "npx_create", "npx_destroy", "npx_last_error", "npx_device_count", "npx_set_params",
    "npx_set_particles", "npx_get_particles", "npx_num_particles", "npx_step", "npx_step_n",
    "npx_synchronize", "npx_debug_grid", "npx_debug_candidates", "npx_grid_query",
    "npx_set_boundary", "npx_compute_boundary_volumes", "npx_boundary_world_positions",
    "npx_apply_boundary_density", "npx_coupling_forces", "npx_batch_create", "npx_scene_add",
    "npx_scene_remove", "npx_scene_get", "npx_batch_step", "npx_scene_bytes",
    "npx_set_kernel_timing", "npx_kernel_times", "npx_kernel_trace", "npx_batch_step_timed", "npx_scene_status",
    "npx_set_tuning", "npx_set_persistent", "npx_set_rank_split", "npx_set_list_policy", "npx_scene_set_boundary", "npx_batch_num_bodies", "npx_batch_coupling",
    "npx_sample_boundaries", "npx_grid_query_noalloc",
)


SHAPE_BOX, SHAPE_SPHERE, SHAPE_PLANE = 0, 1, 2


class NpxShape(C.Structure):
    """npx_shape: the CollisionShape fields the boundary samplers read."""
    _fields_ = [("type", C.c_int32), ("body", C.c_int32), ("half_extents", C.c_float * 3), ("radius", C.c_float),
                ("normal", C.c_float * 3), ("offset", C.c_float), ("position", C.c_float * 3),
                ("rotation", C.c_float * 4)]


class NpxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"npx error {code}: {msg}")
        self.code = code


class DomainError(NpxError):
    pass


class NpxParams(C.Structure):
    """npx_pbf_params, same layout as PBFSettings (pbf_solver.h:13-38)."""
    _fields_ = [
        ("rest_density", C.c_float), ("kernel_radius", C.c_float),
        ("particle_radius", C.c_float), ("solver_iterations", C.c_int32),
        ("epsilon", C.c_float), ("xsph_viscosity", C.c_float),
        ("vorticity_epsilon", C.c_float), ("corr_delta_q", C.c_float),
        ("corr_k", C.c_float), ("corr_n", C.c_int32),
        ("use_domain_bounds", C.c_int32),
        ("domain_lower", C.c_float * 3), ("domain_upper", C.c_float * 3),
    ]


def to_params(p) -> NpxParams:
    o = NpxParams()
    o.rest_density = p.rest_density
    o.kernel_radius = p.kernel_radius
    o.particle_radius = p.particle_radius
    o.solver_iterations = int(p.solver_iterations)
    o.epsilon = p.epsilon
    o.xsph_viscosity = p.xsph_viscosity
    o.vorticity_epsilon = getattr(p, "vorticity_epsilon", 0.0)
    o.corr_delta_q = getattr(p, "corr_delta_q", 0.3)
    o.corr_k = getattr(p, "corr_k", 0.001)
    o.corr_n = int(getattr(p, "corr_n", 4))
    o.use_domain_bounds = 1 if p.use_domain_bounds else 0
    for d in range(3):
        o.domain_lower[d] = p.domain_lower[d]
        o.domain_upper[d] = p.domain_upper[d]
    return o


_lib = None


def load() -> C.CDLL:
    """Loads libnpx_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m novaphy_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, f32 = C.c_void_p, C.c_int32, C.c_int64, C.c_float
    P = C.POINTER
    sig = {
        "npx_create": (C.c_int, [P(NpxParams), i32, C.c_uint32, P(vp)]),
        "npx_destroy": (C.c_int, [vp]),
        "npx_last_error": (C.c_char_p, [vp]),
        "npx_device_count": (i32, []),
        "npx_set_params": (C.c_int, [vp, P(NpxParams)]),
        "npx_set_particles": (C.c_int, [vp, vp, vp, i32]),
        "npx_get_particles": (C.c_int, [vp, vp, vp, vp, vp]),
        "npx_num_particles": (i32, [vp]),
        "npx_step": (C.c_int, [vp, f32, vp, f32]),
        "npx_step_n": (C.c_int, [vp, f32, vp, f32, i32]),
        "npx_synchronize": (C.c_int, [vp]),
        "npx_debug_grid": (C.c_int, [vp, vp, vp, vp]),
        "npx_debug_candidates": (C.c_int, [vp, vp, i64, P(i64)]),
        "npx_grid_query": (C.c_int, [vp, i32, f32, vp, i32, vp, vp, i64, P(i64)]),
        "npx_grid_query_noalloc": (C.c_int, [vp, i32, f32, vp, i32, vp, vp, i64, P(i64)]),
        "npx_set_boundary": (C.c_int, [vp, vp, vp, vp, i32, i32]),
        "npx_compute_boundary_volumes": (C.c_int, [vp, vp, i32, vp]),
        "npx_boundary_world_positions": (C.c_int, [vp, vp, i32, vp]),
        "npx_apply_boundary_density": (C.c_int, [vp]),
        "npx_coupling_forces": (C.c_int, [vp, vp, vp, i32, vp, vp]),
        "npx_batch_create": (C.c_int, [i32, i32, i32, C.c_uint32, P(vp)]),
        "npx_scene_add": (C.c_int, [vp, P(NpxParams), f32, vp, f32, vp, vp, i32, P(i32)]),
        "npx_scene_remove": (C.c_int, [vp, i32]),
        "npx_scene_get": (C.c_int, [vp, i32, vp, vp, vp, vp]),
        "npx_batch_step": (C.c_int, [vp, i32]),
        "npx_scene_bytes": (i64, [P(NpxParams), i32]),
        "npx_set_kernel_timing": (C.c_int, [vp, i32]),
        "npx_kernel_times": (C.c_int, [vp, C.c_char_p, i32, vp, vp, i32, P(i32)]),
        "npx_kernel_trace": (C.c_int, [vp, C.c_char_p, i32, vp, vp, i32, P(i32)]),
        "npx_batch_step_timed": (C.c_int, [vp, i32, vp]),
        "npx_scene_status": (C.c_int, [vp, vp, i32]),
        "npx_set_tuning": (C.c_int, [vp, i32, i32, f32]),
        "npx_set_persistent": (C.c_int, [vp, i32]),
        "npx_set_rank_split": (C.c_int, [vp, i32]),
        "npx_set_list_policy": (C.c_int, [vp, i32]),
        "npx_scene_set_boundary": (C.c_int, [vp, i32, vp, vp, vp, i32, i32, vp]),
        "npx_batch_num_bodies": (i32, [vp]),
        "npx_batch_coupling": (C.c_int, [vp, vp, i32, vp]),
        "npx_sample_boundaries": (C.c_int, [vp, i32, f32, f32, vp, vp, i64, P(i64)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def device_count() -> int:
    return int(load().npx_device_count())


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f32(a, shape_last=None) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a


def _check(rc: int, ctx=None):
    if rc == NPX_OK:
        return
    msg = load().npx_last_error(ctx).decode(errors="replace")
    if rc == NPX_EDOMAIN:
        raise DomainError(rc, msg)
    if rc == NPX_EINVAL:
        raise ValueError(msg)
    raise NpxError(rc, msg)


class Context:
    """One single-scene device context (PBFSolver + its particle state)."""

    def __init__(self, params, capacity: int = 1, flags: int = 0):
        self._L = load()
        self._prm = to_params(params)
        h = C.c_void_p()
        _check(self._L.npx_create(C.byref(self._prm), int(capacity), int(flags), C.byref(h)))
        self._h = h
        self.n = 0

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            self._L.npx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _c(self, rc):
        _check(rc, self._h)

    def set_params(self, params):
        self._prm = to_params(params)
        self._c(self._L.npx_set_params(self._h, C.byref(self._prm)))

    def set_particles(self, pos, vel=None):
        pos = _f32(pos).reshape(-1, 3)
        vel = None if vel is None else _f32(vel).reshape(-1, 3)
        self.n = pos.shape[0]
        self._c(self._L.npx_set_particles(self._h, _p(pos), _p(vel), self.n))

    def get_particles(self):
        n = self.n
        pos = np.zeros((n, 3), np.float32)
        vel = np.zeros((n, 3), np.float32)
        rho = np.zeros(n, np.float32)
        lam = np.zeros(n, np.float32)
        self._c(self._L.npx_get_particles(self._h, _p(pos), _p(vel), _p(rho), _p(lam)))
        return pos, vel, rho, lam

    def step(self, dt, gravity, mass, nsteps: int = 1):
        g = _f32(gravity).reshape(3)
        self._c(self._L.npx_step_n(self._h, C.c_float(dt), _p(g), C.c_float(mass), int(nsteps)))

    def synchronize(self):
        self._c(self._L.npx_synchronize(self._h))

    def debug_grid(self):
        n = self.n
        keys = np.zeros(n, np.uint64)
        cand = np.zeros(n, np.int32)
        nbr = np.zeros(n, np.int32)
        self._c(self._L.npx_debug_grid(self._h, _p(keys), _p(cand), _p(nbr)))
        return keys, cand, nbr

    def debug_candidates(self):
        tot = C.c_int64()
        self._c(self._L.npx_debug_candidates(self._h, None, 0, C.byref(tot)))
        out = np.zeros(max(tot.value, 1), np.int32)
        self._c(self._L.npx_debug_candidates(self._h, _p(out), tot.value, C.byref(tot)))
        return out[:tot.value]

    # -- Akinci boundary
    def set_boundary(self, local, body, psi=None, nbodies: int = 0):
        local = _f32(local).reshape(-1, 3)
        body = np.ascontiguousarray(body, np.int32)
        psi = None if psi is None else _f32(psi)
        self.nb = local.shape[0]
        self.nbodies = int(nbodies)
        self._c(self._L.npx_set_boundary(self._h, _p(local), _p(body), _p(psi), self.nb, self.nbodies))

    def compute_boundary_volumes(self, poses7):
        poses7 = _f32(poses7).reshape(-1)
        psi = np.zeros(self.nb, np.float32)
        self._c(self._L.npx_compute_boundary_volumes(self._h, _p(poses7), self.nbodies, _p(psi)))
        return psi

    def boundary_world_positions(self, poses7, read_back: bool = False):
        poses7 = _f32(poses7).reshape(-1)
        out = np.zeros((self.nb, 3), np.float32) if read_back else None
        self._c(self._L.npx_boundary_world_positions(self._h, _p(poses7), self.nbodies, _p(out)))
        return out

    def apply_boundary_density(self):
        self._c(self._L.npx_apply_boundary_density(self._h))

    def coupling_forces(self, poses7, is_static):
        poses7 = _f32(poses7).reshape(-1)
        st = np.ascontiguousarray(is_static, np.uint8)
        f = np.zeros((self.nbodies, 3), np.float32)
        t = np.zeros((self.nbodies, 3), np.float32)
        self._c(self._L.npx_coupling_forces(self._h, _p(poses7), _p(st), self.nbodies, _p(f), _p(t)))
        return f, t

    # -- measurement / tuning
    def set_kernel_timing(self, on: bool):
        self._c(self._L.npx_set_kernel_timing(self._h, 1 if on else 0))

    def kernel_times(self):
        return _kernel_times(self._L, self._h)

    def kernel_trace(self):
        """Timed launches in order: [(name, start_ms, dur_ms)]."""
        cnt = C.c_int32()
        _check(self._L.npx_kernel_trace(self._h, None, 0, None, None, 0, C.byref(cnt)), self._h)
        n = cnt.value
        names = C.create_string_buffer(max(64, 48 * n))
        st = np.zeros(max(n, 1), np.float32)
        du = np.zeros(max(n, 1), np.float32)
        _check(self._L.npx_kernel_trace(self._h, names, len(names), _p(st), _p(du), n, C.byref(cnt)), self._h)
        keys = names.value.decode().split("\n")[:n]
        return [(k, float(st[i]), float(du[i])) for i, k in enumerate(keys)]

    def set_tuning(self, block_size: int, xcd_remap: int = 1, skin_fraction: float = 0.1, persistent: int = 0):
        """Launch tunables (results are bitwise independent of them); persistent > 0
        caps each neighbour pass at that many per-lane workgroups (a multiple of 8),
        each looping over particle blocks."""
        self._c(self._L.npx_set_tuning(self._h, int(block_size), int(xcd_remap), C.c_float(skin_fraction)))
        self._c(self._L.npx_set_persistent(self._h, int(persistent)))

    def set_rank_split(self, huge_levels: int = 6):
        """Cell classes spanning more than 2**huge_levels 256-slot blocks rank
        by original index in index space instead of by merge launches (results
        are bitwise independent of it)."""
        self._c(self._L.npx_set_rank_split(self._h, int(huge_levels)))

    def set_list_policy(self, policy: int = 0):
        """0: no skin lists (every testing pass sweeps inline); 1: skin lists
        in steps where the previous step's sampled list bounds mostly held
        (results are bitwise independent of it)."""
        self._c(self._L.npx_set_list_policy(self._h, int(policy)))


def _kernel_times(L, h):
    cnt = C.c_int32()
    names = C.create_string_buffer(8192)
    ms = np.zeros(128, np.float32)
    nl = np.zeros(128, np.int32)
    _check(L.npx_kernel_times(h, names, 8192, _p(ms), _p(nl), 128, C.byref(cnt)), h)
    keys = names.value.decode().split("\n")[:cnt.value]
    return {k: (float(ms[i]), int(nl[i])) for i, k in enumerate(keys)}


def grid_query(points, cell_size: float, queries):
    """SpatialHashGrid(cell).build(points); query_neighbors(q) for each q:
    returns (counts, concatenated candidate lists)."""
    L = load()
    pts = _f32(points).reshape(-1, 3)
    q = _f32(queries).reshape(-1, 3)
    tot = C.c_int64()
    counts = np.zeros(q.shape[0], np.int32)
    _check(L.npx_grid_query(_p(pts), pts.shape[0], C.c_float(cell_size), _p(q), q.shape[0], _p(counts),
                            None, 0, C.byref(tot)))
    out = np.zeros(max(tot.value, 1), np.int32)
    _check(L.npx_grid_query(_p(pts), pts.shape[0], C.c_float(cell_size), _p(q), q.shape[0], _p(counts),
                            _p(out), tot.value, C.byref(tot)))
    return counts, out[:tot.value]


def shapes_from_model(model):
    """npx_shape records of a Model's collision shapes (ModelBuilder output):
    type, body, box half extents, sphere radius, plane normal / offset and
    the shape's local transform."""
    kind = {"Box": SHAPE_BOX, "Sphere": SHAPE_SPHERE, "Plane": SHAPE_PLANE}
    out = (NpxShape * max(len(model.shapes), 1))()
    for k, sh in enumerate(model.shapes):
        r = out[k]
        r.type = kind[str(sh.type).split(".")[-1]]
        r.body = int(sh.body_index)
        r.half_extents[:] = [float(v) for v in np.asarray(sh.half_extents, np.float32)]
        r.radius = float(sh.radius)
        r.normal[:] = [float(v) for v in np.asarray(sh.normal, np.float32)]
        r.offset = float(sh.offset)
        t = sh.local_transform
        r.position[:] = [float(v) for v in np.asarray(t.position, np.float32)]
        x, y, z, w = np.asarray(t.rotation, np.float32)  # xyzw
        r.rotation[:] = [float(w), float(x), float(y), float(z)]
    return out, len(model.shapes)


def sample_boundaries(shapes, nshapes: int, spacing: float, plane_extent: float):
    """Boundary particles of the shapes on the device (npx_sample_boundaries):
    returns (local positions (n, 3) float32, owning body (n,) int32)."""
    L = load()
    cnt = C.c_int64()
    _check(L.npx_sample_boundaries(shapes, int(nshapes), C.c_float(spacing), C.c_float(plane_extent), None, None, 0,
                                   C.byref(cnt)))
    n = cnt.value
    local = np.zeros((max(n, 1), 3), np.float32)
    body = np.zeros(max(n, 1), np.int32)
    _check(L.npx_sample_boundaries(shapes, int(nshapes), C.c_float(spacing), C.c_float(plane_extent), _p(local),
                                   _p(body), n, C.byref(cnt)))
    return local[:n], body[:n]


class Batch:
    """Many independent scenes advanced together (npx_batch_*)."""

    def __init__(self, slot_capacity: int, cell_capacity: int, max_scenes: int, flags: int = 0):
        self._L = load()
        h = C.c_void_p()
        _check(self._L.npx_batch_create(int(slot_capacity), int(cell_capacity), int(max_scenes),
                                        int(flags), C.byref(h)))
        self._h = h
        self._n = {}

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            self._L.npx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc):
        _check(rc, self._h)

    def add(self, params, dt, gravity, mass, pos, vel=None) -> int:
        prm = to_params(params)
        pos = _f32(pos).reshape(-1, 3)
        vel = None if vel is None else _f32(vel).reshape(-1, 3)
        g = _f32(gravity).reshape(3)
        sid = C.c_int32()
        self._c(self._L.npx_scene_add(self._h, C.byref(prm), C.c_float(dt), _p(g), C.c_float(mass),
                                      _p(pos), _p(vel), pos.shape[0], C.byref(sid)))
        self._n[sid.value] = pos.shape[0]
        return sid.value

    def set_boundary(self, sid: int, local, body, psi, nbodies: int, is_static):
        """Akinci boundary of scene sid (npx_scene_set_boundary)."""
        local = _f32(local).reshape(-1, 3)
        body = np.ascontiguousarray(body, np.int32).reshape(-1)
        psi = _f32(psi).reshape(-1)
        st = np.ascontiguousarray(is_static, np.uint8).reshape(-1)
        if not (local.shape[0] == body.shape[0] == psi.shape[0]) or st.shape[0] != nbodies:
            raise ValueError("boundary array sizes do not match")
        self._c(self._L.npx_scene_set_boundary(self._h, int(sid), _p(local), _p(body), _p(psi), local.shape[0],
                                               int(nbodies), _p(st)))

    def num_bodies(self) -> int:
        n = int(self._L.npx_batch_num_bodies(self._h))
        if n < 0:
            self._c(-1)
        return n

    def coupling(self, poses7) -> np.ndarray:
        """Boundary density and coupling forces of every scene with a boundary
        after a step; returns (bodies, 6) force + torque (npx_batch_coupling)."""
        poses7 = _f32(poses7).reshape(-1, 7)
        out = np.zeros((max(poses7.shape[0], 1), 6), np.float32)
        self._c(self._L.npx_batch_coupling(self._h, _p(poses7), poses7.shape[0], _p(out)))
        return out[:poses7.shape[0]]

    def remove(self, sid: int):
        self._c(self._L.npx_scene_remove(self._h, int(sid)))
        self._n.pop(sid, None)

    def get(self, sid: int):
        n = self._n[sid]
        pos = np.zeros((n, 3), np.float32)
        vel = np.zeros((n, 3), np.float32)
        rho = np.zeros(n, np.float32)
        lam = np.zeros(n, np.float32)
        self._c(self._L.npx_scene_get(self._h, int(sid), _p(pos), _p(vel), _p(rho), _p(lam)))
        return pos, vel, rho, lam

    def step(self, nsteps: int = 1):
        self._c(self._L.npx_batch_step(self._h, int(nsteps)))

    def step_timed(self, nsteps: int = 1) -> np.ndarray:
        """Steps with device timestamps; returns per-step milliseconds."""
        out = np.zeros(max(int(nsteps), 1), np.float32)
        self._c(self._L.npx_batch_step_timed(self._h, int(nsteps), _p(out)))
        return out[:int(nsteps)]

    def status(self, max_scenes: int) -> np.ndarray:
        """Per-scene NPX_STATUS_* bits since the last call (cleared on read)."""
        out = np.zeros(int(max_scenes), np.int32)
        self._c(self._L.npx_scene_status(self._h, _p(out), int(max_scenes)))
        return out

    def synchronize(self):
        self._c(self._L.npx_synchronize(self._h))

    def set_kernel_timing(self, on: bool):
        self._c(self._L.npx_set_kernel_timing(self._h, 1 if on else 0))

    def kernel_times(self):
        return _kernel_times(self._L, self._h)

    def set_tuning(self, block_size: int, xcd_remap: int = 1, skin_fraction: float = 0.1, persistent: int = 0):
        """Launch tunables (results are bitwise independent of them); persistent > 0
        caps each neighbour pass at that many per-lane workgroups (a multiple of 8),
        each looping over particle blocks."""
        self._c(self._L.npx_set_tuning(self._h, int(block_size), int(xcd_remap), C.c_float(skin_fraction)))
        self._c(self._L.npx_set_persistent(self._h, int(persistent)))

    def set_rank_split(self, huge_levels: int = 6):
        """Cell classes spanning more than 2**huge_levels 256-slot blocks rank
        by original index in index space instead of by merge launches (results
        are bitwise independent of it)."""
        self._c(self._L.npx_set_rank_split(self._h, int(huge_levels)))

    def set_list_policy(self, policy: int = 0):
        """0: no skin lists (every testing pass sweeps inline); 1: skin lists
        in steps where the previous step's sampled list bounds mostly held
        (results are bitwise independent of it)."""
        self._c(self._L.npx_set_list_policy(self._h, int(policy)))


def scene_bytes(params, n: int) -> int:
    prm = to_params(params)
    return int(load().npx_scene_bytes(C.byref(prm), int(n)))
