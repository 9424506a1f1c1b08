"""ctypes wrapper around liboracle (oracle/pbf_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpbf_oracle.so")


class OrcParams(C.Structure):
    _fields_ = [
        ("rest_density", C.c_float), ("kernel_radius", C.c_float),
        ("particle_radius", C.c_float), ("solver_iterations", C.c_int32),
        ("epsilon", C.c_float), ("xsph_viscosity", C.c_float),
        ("vorticity_epsilon", C.c_float), ("corr_delta_q", C.c_float),
        ("corr_k", C.c_float), ("corr_n", C.c_int32),
        ("use_domain_bounds", C.c_int32),
        ("domain_lower", C.c_float * 3), ("domain_upper", C.c_float * 3),
    ]


class OrcShape(C.Structure):
    _fields_ = [("type", C.c_int32), ("body", C.c_int32), ("a", C.c_float * 4),
                ("local_pos", C.c_float * 3), ("local_rot", C.c_float * 4)]


class OrcRBody(C.Structure):
    _fields_ = [("mass", C.c_float), ("inertia", C.c_float * 9)]


class OrcRShape(C.Structure):
    _fields_ = [("type", C.c_int32), ("body", C.c_int32), ("half", C.c_float * 3), ("radius", C.c_float),
                ("normal", C.c_float * 3), ("offset", C.c_float), ("local_pos", C.c_float * 3),
                ("local_rot", C.c_float * 4), ("friction", C.c_float), ("restitution", C.c_float)]


class OrcRSettings(C.Structure):
    _fields_ = [("velocity_iterations", C.c_int32), ("baumgarte", C.c_float), ("slop", C.c_float),
                ("warm_starting", C.c_int32)]


class OrcContact(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("normal", C.c_float * 3), ("penetration", C.c_float),
                ("body_a", C.c_int32), ("body_b", C.c_int32), ("friction", C.c_float), ("restitution", C.c_float),
                ("acc_n", C.c_float), ("acc_t1", C.c_float), ("acc_t2", C.c_float)]


_SOURCES = ("pbf_oracle.cpp", "rigid_oracle.cpp")


def build(force: bool = False) -> str:
    newest = max(os.path.getmtime(os.path.join(HERE, f)) for f in _SOURCES)
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < newest:
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
        u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
        vp = C.c_void_p
        L.orc_poly6.restype = C.c_float
        L.orc_poly6.argtypes = [C.c_float, C.c_float]
        L.orc_poly6_f64.restype = C.c_double
        L.orc_poly6_f64.argtypes = [C.c_double, C.c_double]
        L.orc_spiky_grad.argtypes = [f32p, C.c_float, f32p]
        L.orc_spiky.restype = C.c_float
        L.orc_spiky.argtypes = [C.c_float, C.c_float]
        L.orc_cubic_spline.restype = C.c_float
        L.orc_cubic_spline.argtypes = [C.c_float, C.c_float]
        L.orc_hash_cell.restype = C.c_uint64
        L.orc_hash_cell.argtypes = [C.c_int, C.c_int, C.c_int]
        L.orc_generate_fluid_block.restype = C.c_int32
        L.orc_generate_fluid_block.argtypes = [f32p, f32p, C.c_float, vp]
        L.orc_grid_debug.restype = C.c_int64
        L.orc_grid_debug.argtypes = [C.c_int32, f32p, C.c_float, C.c_float, vp, vp, vp, vp,
                                     C.c_int64]
        L.orc_grid_query.restype = C.c_int64
        L.orc_grid_query.argtypes = [C.c_int32, f32p, C.c_float, C.c_int32, f32p, vp, vp,
                                     C.c_int64]
        L.orc_grid_all_pairs.restype = C.c_int64
        L.orc_grid_all_pairs.argtypes = [C.c_int32, f32p, C.c_float, C.c_float, vp, C.c_int64]
        L.orc_pbf_run_f32.restype = C.c_int32
        L.orc_pbf_run_f32.argtypes = [C.POINTER(OrcParams), C.c_int32, f32p, f32p, f32p, f32p,
                                      C.c_float, f32p, C.c_float, C.c_int32]
        L.orc_pbf_run_f64.restype = C.c_int32
        L.orc_pbf_run_f64.argtypes = [C.POINTER(OrcParams), C.c_int32, f64p, f64p, f64p, f64p,
                                      C.c_double, f64p, C.c_double, C.c_int32]
        L.orc_sim_create.restype = vp
        L.orc_sim_create.argtypes = [C.POINTER(OrcParams), C.c_int32, f32p, f32p]
        L.orc_sim_step.restype = C.c_int32
        L.orc_sim_step.argtypes = [vp, C.c_float, f32p, C.c_float, C.c_int32]
        L.orc_sim_get.argtypes = [vp, vp, vp, vp, vp]
        L.orc_sim_destroy.argtypes = [vp]
        L.orc_sample_boundaries.restype = C.c_int32
        L.orc_sample_boundaries.argtypes = [C.c_int32, C.POINTER(OrcShape), C.c_float,
                                            C.c_float, vp, vp]
        L.orc_boundary_world_positions.restype = C.c_int32
        L.orc_boundary_world_positions.argtypes = [C.c_int32, f32p, i32p, C.c_int32, f32p, f32p]
        L.orc_boundary_volumes.restype = C.c_int32
        L.orc_boundary_volumes.argtypes = [C.c_int32, f32p, C.c_float, f32p]
        L.orc_boundary_volumes_f64.restype = C.c_int32
        L.orc_boundary_volumes_f64.argtypes = [C.c_int32, f64p, C.c_double, f64p]
        L.orc_boundary_density.restype = C.c_int32
        L.orc_boundary_density.argtypes = [C.c_int32, f32p, f32p, C.c_int32, f32p, f32p,
                                           C.c_float, C.c_float]
        L.orc_boundary_density_f64.restype = C.c_int32
        L.orc_boundary_density_f64.argtypes = [C.c_int32, f64p, f64p, C.c_int32, f64p, f64p,
                                               C.c_double, C.c_double]
        L.orc_coupling.restype = C.c_int32
        L.orc_coupling.argtypes = [C.c_int32, f32p, f32p, C.c_int32, f32p, f32p, i32p, C.c_int32,
                                   f32p, u8p, C.c_float, C.c_float, C.c_float, f32p, f32p, vp]
        L.orc_coupling_f64.restype = C.c_int32
        L.orc_coupling_f64.argtypes = [C.c_int32, f64p, f64p, C.c_int32, f64p, f64p, i32p,
                                       C.c_int32, f64p, u8p, C.c_double, C.c_double, C.c_double,
                                       f64p, f64p]
        L.orc_collide_shapes.restype = C.c_int32
        L.orc_collide_shapes.argtypes = [C.POINTER(OrcRShape), f32p, C.POINTER(OrcRShape), f32p, vp, C.c_int32]
        L.orc_sap_pairs.restype = C.c_int32
        L.orc_sap_pairs.argtypes = [C.c_int32, f32p, f32p, u8p, i32p, C.c_int32]
        L.orc_rigid_step.restype = C.c_int32
        L.orc_rigid_step.argtypes = [vp, C.c_int32, vp, C.c_int32, C.POINTER(OrcRSettings), f32p, C.c_float, f32p,
                                     f32p, f32p, f32p, f32p, vp, C.c_int32, C.POINTER(C.c_int32),
                                     C.POINTER(C.c_int32)]
        L.orc_vorticity_confinement.restype = C.c_int32
        L.orc_vorticity_confinement.argtypes = [C.POINTER(OrcParams), C.c_int32, f32p, f32p, f32p, C.c_float,
                                                C.c_float, f32p, f32p, f32p, f32p]
        L.orc_vorticity_omega.restype = C.c_int32
        L.orc_vorticity_omega.argtypes = [C.POINTER(OrcParams), C.c_int32, f32p, f32p, f32p,
                                          C.c_float, f32p]
        _ = u64p
        _lib = L
    return _lib


def params(p) -> OrcParams:
    """OrcParams from any object with PBFSettings field names."""
    o = OrcParams()
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


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


# ----------------------------------------------------------------- kernels
def poly6(r_sq, h):
    return lib().orc_poly6(r_sq, h)


def poly6_f64(r_sq, h):
    return lib().orc_poly6_f64(r_sq, h)


def spiky_grad(r, h):
    out = np.zeros(3, np.float32)
    lib().orc_spiky_grad(np.ascontiguousarray(r, np.float32), h, out)
    return out


def spiky(r_len, h):
    return lib().orc_spiky(r_len, h)


def cubic_spline(r_len, h):
    return lib().orc_cubic_spline(r_len, h)


def hash_cell(cx, cy, cz) -> int:
    return int(lib().orc_hash_cell(cx, cy, cz))


def generate_fluid_block(lower, upper, spacing):
    lo = np.ascontiguousarray(lower, np.float32)
    hi = np.ascontiguousarray(upper, np.float32)
    n = lib().orc_generate_fluid_block(lo, hi, spacing, None)
    out = np.zeros((n, 3), np.float32)
    lib().orc_generate_fluid_block(lo, hi, spacing, _ptr(out))
    return out


# ----------------------------------------------------------------- grid
def grid_debug(pos, cell, radius, with_lists=False):
    """keys, cand_cnt, nbr_cnt (and flattened candidate lists) of one build."""
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    n = pos.shape[0]
    keys = np.zeros(n, np.uint64)
    cc = np.zeros(n, np.int32)
    nc = np.zeros(n, np.int32)
    total = lib().orc_grid_debug(n, pos, cell, radius, _ptr(keys), _ptr(cc), _ptr(nc), None, 0)
    lists = None
    if with_lists:
        lists = np.zeros(total, np.int32)
        lib().orc_grid_debug(n, pos, cell, radius, None, None, None, _ptr(lists), total)
    return keys, cc, nc, lists


def grid_query(pos, cell, queries):
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    q = np.ascontiguousarray(queries, np.float32).reshape(-1, 3)
    counts = np.zeros(q.shape[0], np.int32)
    total = lib().orc_grid_query(pos.shape[0], pos, cell, q.shape[0], q, _ptr(counts), None, 0)
    lst = np.zeros(total, np.int32)
    lib().orc_grid_query(pos.shape[0], pos, cell, q.shape[0], q, None, _ptr(lst), total)
    return counts, lst


def grid_all_pairs(pos, cell, radius):
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    total = lib().orc_grid_all_pairs(pos.shape[0], pos, cell, radius, None, 0)
    pairs = np.zeros((total, 2), np.int32)
    lib().orc_grid_all_pairs(pos.shape[0], pos, cell, radius, _ptr(pairs), total)
    return pairs


# ----------------------------------------------------------------- PBF
def pbf_run(prm, pos, vel, dt, gravity, mass, nsteps=1):
    """nsteps of the FP32 restatement; returns (pos, vel, rho, lambda)."""
    pos = np.array(pos, np.float32, copy=True).reshape(-1, 3)
    vel = np.array(vel, np.float32, copy=True).reshape(-1, 3)
    n = pos.shape[0]
    rho = np.zeros(n, np.float32)
    lam = np.zeros(n, np.float32)
    op = params(prm)
    lib().orc_pbf_run_f32(C.byref(op), n, pos, vel, rho, lam, np.float32(dt),
                          np.ascontiguousarray(gravity, np.float32), np.float32(mass), nsteps)
    return pos, vel, rho, lam


def pbf_run_f64(prm, pos, vel, dt, gravity, mass, nsteps=1):
    """The same steps in FP64 (inputs widened exactly from float32)."""
    pos = np.array(pos, np.float64, copy=True).reshape(-1, 3)
    vel = np.array(vel, np.float64, copy=True).reshape(-1, 3)
    n = pos.shape[0]
    rho = np.zeros(n, np.float64)
    lam = np.zeros(n, np.float64)
    op = params(prm)
    lib().orc_pbf_run_f64(C.byref(op), n, pos, vel, rho, lam, float(np.float32(dt)),
                          np.ascontiguousarray(np.float32(gravity), np.float64),
                          float(np.float32(mass)), nsteps)
    return pos, vel, rho, lam


class Sim:
    """Persistent FP32 restatement state (for timing)."""

    def __init__(self, prm, pos, vel):
        self._op = params(prm)
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, np.float32).reshape(-1, 3)
        self.n = pos.shape[0]
        self._h = lib().orc_sim_create(C.byref(self._op), self.n, pos, vel)

    def step(self, dt, gravity, mass, nsteps=1):
        lib().orc_sim_step(self._h, np.float32(dt), np.ascontiguousarray(gravity, np.float32),
                           np.float32(mass), nsteps)

    def get(self):
        pos = np.zeros((self.n, 3), np.float32)
        vel = np.zeros((self.n, 3), np.float32)
        rho = np.zeros(self.n, np.float32)
        lam = np.zeros(self.n, np.float32)
        lib().orc_sim_get(self._h, _ptr(pos), _ptr(vel), _ptr(rho), _ptr(lam))
        return pos, vel, rho, lam

    def close(self):
        if self._h:
            lib().orc_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ----------------------------------------------------------------- boundary
SHAPE_BOX, SHAPE_SPHERE, SHAPE_PLANE = 0, 1, 2


def make_shape(kind, body, a, local_pos=(0, 0, 0), local_rot=(1, 0, 0, 0)):
    s = OrcShape()
    s.type = kind
    s.body = body
    for i, v in enumerate(a):
        s.a[i] = v
    for i in range(3):
        s.local_pos[i] = local_pos[i]
    for i in range(4):
        s.local_rot[i] = local_rot[i]
    return s


def sample_boundaries(shapes, spacing, plane_extent=1.0):
    arr = (OrcShape * len(shapes))(*shapes)
    n = lib().orc_sample_boundaries(len(shapes), arr, spacing, plane_extent, None, None)
    local = np.zeros((n, 3), np.float32)
    body = np.zeros(n, np.int32)
    lib().orc_sample_boundaries(len(shapes), arr, spacing, plane_extent, _ptr(local), _ptr(body))
    return local, body


def boundary_world_positions(local, body, poses7):
    local = np.ascontiguousarray(local, np.float32).reshape(-1, 3)
    body = np.ascontiguousarray(body, np.int32)
    poses7 = np.ascontiguousarray(poses7, np.float32).reshape(-1, 7)
    out = np.zeros_like(local)
    lib().orc_boundary_world_positions(local.shape[0], local, body, poses7.shape[0],
                                       poses7.reshape(-1) if poses7.size else np.zeros(7, np.float32),
                                       out)
    return out


def boundary_volumes(wpos, h):
    wpos = np.ascontiguousarray(wpos, np.float32).reshape(-1, 3)
    psi = np.zeros(wpos.shape[0], np.float32)
    lib().orc_boundary_volumes(wpos.shape[0], wpos, h, psi)
    return psi


def boundary_volumes_f64(wpos, h):
    wpos = np.ascontiguousarray(wpos, np.float64).reshape(-1, 3)
    psi = np.zeros(wpos.shape[0], np.float64)
    lib().orc_boundary_volumes_f64(wpos.shape[0], wpos, float(np.float32(h)), psi)
    return psi


def boundary_density(pos, rho, bpos, psi, h, rho0):
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    rho = np.array(rho, np.float32, copy=True)
    bpos = np.ascontiguousarray(bpos, np.float32).reshape(-1, 3)
    psi = np.ascontiguousarray(psi, np.float32)
    lib().orc_boundary_density(pos.shape[0], pos, rho, bpos.shape[0], bpos, psi, h, rho0)
    return rho


def coupling(pos, rho, bpos, psi, body, poses7, is_static, h, rho0, mass, per_particle=False):
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    rho = np.ascontiguousarray(rho, np.float32)
    bpos = np.ascontiguousarray(bpos, np.float32).reshape(-1, 3)
    psi = np.ascontiguousarray(psi, np.float32)
    body = np.ascontiguousarray(body, np.int32)
    poses7 = np.ascontiguousarray(poses7, np.float32).reshape(-1, 7)
    nb = poses7.shape[0]
    st = np.ascontiguousarray(is_static, np.uint8)
    f = np.zeros((nb, 3), np.float32)
    t = np.zeros((nb, 3), np.float32)
    pp = np.zeros((bpos.shape[0], 3), np.float32) if per_particle else None
    lib().orc_coupling(pos.shape[0], pos, rho, bpos.shape[0], bpos, psi, body, nb,
                       poses7.reshape(-1), st, h, rho0, mass, f.reshape(-1), t.reshape(-1), _ptr(pp))
    return (f, t, pp) if per_particle else (f, t)


def coupling_f64(pos, rho, bpos, psi, body, poses7, is_static, h, rho0, mass):
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    rho = np.ascontiguousarray(rho, np.float64)
    bpos = np.ascontiguousarray(bpos, np.float64).reshape(-1, 3)
    psi = np.ascontiguousarray(psi, np.float64)
    body = np.ascontiguousarray(body, np.int32)
    poses7 = np.ascontiguousarray(poses7, np.float64).reshape(-1, 7)
    nb = poses7.shape[0]
    st = np.ascontiguousarray(is_static, np.uint8)
    f = np.zeros((nb, 3), np.float64)
    t = np.zeros((nb, 3), np.float64)
    lib().orc_coupling_f64(pos.shape[0], pos, rho, bpos.shape[0], bpos, psi, body, nb,
                           poses7.reshape(-1), st, float(np.float32(h)), float(np.float32(rho0)),
                           float(np.float32(mass)), f.reshape(-1), t.reshape(-1))
    return f, t


def vorticity_omega(prm, pos, vel, rho, mass):
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    vel = np.ascontiguousarray(vel, np.float32).reshape(-1, 3)
    rho = np.ascontiguousarray(rho, np.float32)
    om = np.zeros_like(pos)
    op = params(prm)
    lib().orc_vorticity_omega(C.byref(op), pos.shape[0], pos, vel, rho, np.float32(mass), om)
    return om


def vorticity_confinement(prm, pos, vel, rho, mass, dt):
    """omega, eta = grad|omega|, confinement force f and v + dt f per particle
    (this project's vorticity confinement, see pbf_oracle.cpp)."""
    op = params(prm)
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    vel = np.ascontiguousarray(vel, np.float32).reshape(-1, 3)
    rho = np.ascontiguousarray(rho, np.float32).reshape(-1)
    n = pos.shape[0]
    outs = [np.zeros((n, 3), np.float32) for _ in range(4)]
    lib().orc_vorticity_confinement(C.byref(op), n, pos, vel, rho, np.float32(mass), np.float32(dt), *outs)
    return tuple(outs)


# ------------------------------------------------------------ rigid bodies
SHAPE_R_BOX, SHAPE_R_SPHERE, SHAPE_R_PLANE = 0, 1, 2


def rigid_shape(kind, body, half=(0.5, 0.5, 0.5), radius=0.5, normal=(0, 1, 0), offset=0.0,
                local_pos=(0, 0, 0), local_rot=(1, 0, 0, 0), friction=0.5, restitution=0.3):
    """CollisionShape fields (shape.h:43-60); local_rot is (w, x, y, z)."""
    s = OrcRShape()
    s.type, s.body = int(kind), int(body)
    s.half[:] = [float(v) for v in half]
    s.radius, s.offset = float(radius), float(offset)
    s.normal[:] = [float(v) for v in normal]
    s.local_pos[:] = [float(v) for v in local_pos]
    s.local_rot[:] = [float(v) for v in local_rot]
    s.friction, s.restitution = float(friction), float(restitution)
    return s


def rigid_body(mass, inertia):
    b = OrcRBody()
    b.mass = float(mass)
    b.inertia[:] = [float(v) for v in np.asarray(inertia, np.float32).reshape(9)]
    return b


def _contacts(buf, n):
    keys = ("position", "normal", "penetration", "body_a", "body_b", "friction", "restitution", "acc_n",
            "acc_t1", "acc_t2")
    out = []
    for c in buf[:n]:
        d = {k: getattr(c, k) for k in keys}
        d["position"] = np.array(d["position"][:], np.float32)
        d["normal"] = np.array(d["normal"][:], np.float32)
        out.append(d)
    return out


def collide_shapes(a, pose_a7, b, pose_b7):
    """Narrowphase of one shape pair (narrowphase.cpp:384-432): list of contact dicts."""
    buf = (OrcContact * 16)()
    n = lib().orc_collide_shapes(C.byref(a), np.ascontiguousarray(pose_a7, np.float32), C.byref(b),
                                 np.ascontiguousarray(pose_b7, np.float32), buf, 16)
    return _contacts(buf, n)


def sap_pairs(lo, hi, static):
    lo = np.ascontiguousarray(lo, np.float32).reshape(-1, 3)
    hi = np.ascontiguousarray(hi, np.float32).reshape(-1, 3)
    st = np.ascontiguousarray(static, np.uint8)
    n = lo.shape[0]
    cap = n * (n - 1) // 2 + 1
    out = np.zeros(2 * cap, np.int32)
    k = lib().orc_sap_pairs(n, lo, hi, st, out, cap)
    return [tuple(int(v) for v in out[2 * i:2 * i + 2]) for i in range(k)]


class RigidWorld:
    """World::step_rigid_pipeline restated (rigid_oracle.cpp): bodies and
    shapes as built by a ModelBuilder, state as float32 arrays."""

    def __init__(self, bodies, shapes, poses7, velocity_iterations=30, baumgarte=0.3, slop=0.005,
                 warm_starting=True, gravity=(0.0, -9.81, 0.0)):
        self.bodies = (OrcRBody * max(len(bodies), 1))(*bodies)
        self.nb = len(bodies)
        self.shapes = (OrcRShape * max(len(shapes), 1))(*shapes)
        self.ns = len(shapes)
        self.settings = OrcRSettings(int(velocity_iterations), float(baumgarte), float(slop), int(warm_starting))
        self.gravity = np.asarray(gravity, np.float32)
        self.pose7 = np.ascontiguousarray(poses7, np.float32).reshape(-1, 7).copy()
        z = np.zeros((self.nb, 3), np.float32)
        self.lin, self.ang, self.force, self.torque = z.copy(), z.copy(), z.copy(), z.copy()
        self.contacts = []
        self.npairs = 0

    def apply_force(self, i, f):
        if 0 <= i < self.nb:
            self.force[i] = (self.force[i] + np.asarray(f, np.float32)).astype(np.float32)

    def apply_torque(self, i, t):
        if 0 <= i < self.nb:
            self.torque[i] = (self.torque[i] + np.asarray(t, np.float32)).astype(np.float32)

    def step(self, dt):
        cap = 64 * max(self.ns, 1)
        buf = (OrcContact * cap)()
        nc, npairs = C.c_int32(), C.c_int32()
        for a in (self.pose7, self.lin, self.ang, self.force, self.torque):
            assert a.flags["C_CONTIGUOUS"] and a.dtype == np.float32
        lib().orc_rigid_step(self.bodies, self.nb, self.shapes, self.ns, C.byref(self.settings), self.gravity,
                             np.float32(dt), self.pose7.reshape(-1), self.lin.reshape(-1), self.ang.reshape(-1),
                             self.force.reshape(-1), self.torque.reshape(-1), buf, cap, C.byref(nc),
                             C.byref(npairs))
        self.contacts = _contacts(buf, min(nc.value, cap))
        self.npairs = npairs.value


def _sqn32(v):
    v = np.asarray(v, np.float32)
    return np.float32(v[0] * v[0] + np.float32(v[1] * v[1] + v[2] * v[2]))


class FluidWorldRun:
    """FluidWorld::step restated (fluid_world.cpp:46-174) on the oracle:
    boundary world positions at the bodies' poses before the step, the PBF
    step, boundary density on the committed positions, the per-body coupling
    force and torque (reduced in boundary-particle order), handed to the
    rigid world as World::apply_force / apply_torque (skipped for static
    bodies and for a zero float32 force and torque), then the rigid step."""

    def __init__(self, prm, pos, vel, mass, local, body, psi, rigid, is_static):
        self.prm, self.mass = prm, float(np.float32(mass))
        self.pos = np.ascontiguousarray(pos, np.float32).copy()
        self.vel = np.ascontiguousarray(vel, np.float32).copy()
        self.rho = np.zeros(self.pos.shape[0], np.float32)
        self.lam = np.zeros(self.pos.shape[0], np.float32)
        self.local = np.ascontiguousarray(local, np.float32).reshape(-1, 3)
        self.body = np.ascontiguousarray(body, np.int32)
        self.psi = np.ascontiguousarray(psi, np.float32)
        self.rigid = rigid
        self.is_static = np.ascontiguousarray(is_static, np.uint8)

    def step(self, dt, gravity):
        h, rho0 = float(np.float32(self.prm.kernel_radius)), float(np.float32(self.prm.rest_density))
        poses = self.rigid.pose7.copy()
        nb = self.local.shape[0]
        bpos = boundary_world_positions(self.local, self.body, poses) if nb else None
        self.pos, self.vel, self.rho, self.lam = pbf_run(self.prm, self.pos, self.vel, dt, gravity, self.mass, 1)
        if nb:
            self.rho = boundary_density(self.pos, self.rho, bpos, self.psi, h, rho0)
            f, t = coupling(self.pos, self.rho, bpos, self.psi, self.body, poses, self.is_static, h, rho0, self.mass)
            for k in range(poses.shape[0]):
                if self.is_static[k] or (_sqn32(f[k]) == 0.0 and _sqn32(t[k]) == 0.0):
                    continue
                self.rigid.apply_force(k, f[k])
                self.rigid.apply_torque(k, t[k])
        self.rigid.step(dt)
