"""Synthetic dam-break scenes used by the tests and by bench.py.

The lattice follows ``generate_fluid_block`` (src/fluid/particle_state.cpp:31-58):
``n_axis = max(1, floor((hi - lo) / s) + 1)``, z outer, y, x inner, positions
``lo + i * s`` in float32.  On top of the lattice every coordinate gets a seeded
jitter that CPU, GPU and numpy reproduce bit for bit:

    u = (splitmix64(seed ^ (3*i + d)) >> 40) * 2**-24          (exact in float32)
    delta = (2u - 1) * 0.001                                    (float32 ops)

splitmix64 is the published mixer of Steele, Lea and Flood (2014).

Scene geometry (all in metres, FP32):
  * dam_break_8k     20x20x20 block at spacing 0.05 with its lower corner at
                     (r, r, r), domain [0,0,0]-[2,1,1], 200 steps, seed 0.
  * dam_break_1m     100x100x100 block, domain [0,0,0]-[10,5,5] (a 100-particle
                     lattice at 0.05 spans 4.95 m, so the box keeps the 2:1:1
                     aspect at 10x5x5 m), vorticity 1e-4, 500 steps, seed 1.
  * coupled_500k     100x50x100 block (y has 50 layers) in [0,0,0]-[10,5,5]
                     plus 64 rigid bodies, seed 2.
  * batched_1024     1024 dam breaks, N log-uniform in [2k, 64k], seed 3.
  * served requests  dam breaks with N log-uniform in [4k, 256k] and 0-8 rigid
                     spheres (r ~ U(0.1, 0.2), density ~ U(300, 2000)) dropped
                     0.3-1.0 m above the block, seed 4.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Tuple

import numpy as np

MASK64 = (1 << 64) - 1


def splitmix64(x: np.ndarray) -> np.ndarray:
    """Vectorised splitmix64 finaliser over uint64 (wrapping arithmetic)."""
    z = (x + np.uint64(0x9E3779B97F4A7C15)).astype(np.uint64)
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)).astype(np.uint64)
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)).astype(np.uint64)
    return z ^ (z >> np.uint64(31))


def jitter(n: int, seed: int, amplitude: float = 0.001) -> np.ndarray:
    """(n, 3) float32 jitter in [-amplitude, amplitude)."""
    idx = np.arange(3 * n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        k = splitmix64(np.uint64(seed) ^ idx)
    u = (k >> np.uint64(40)).astype(np.float32) * np.float32(2.0 ** -24)
    d = (np.float32(2.0) * u - np.float32(1.0)) * np.float32(amplitude)
    return d.astype(np.float32).reshape(n, 3)


def generate_fluid_block(lower, upper, spacing: float) -> np.ndarray:
    """Lattice block, same count rule, order and float32 ops as the reference
    ``generate_fluid_block`` (particle_state.cpp:31-58)."""
    lo = np.asarray(lower, dtype=np.float32)
    hi = np.asarray(upper, dtype=np.float32)
    s = np.float32(spacing)
    if s <= 0:
        return np.zeros((0, 3), np.float32)
    n = [max(1, int(np.floor(np.float32(hi[d] - lo[d]) / s)) + 1) for d in range(3)]
    iz, iy, ix = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    pos = np.empty((n[0] * n[1] * n[2], 3), np.float32)
    pos[:, 0] = lo[0] + ix.ravel().astype(np.float32) * s
    pos[:, 1] = lo[1] + iy.ravel().astype(np.float32) * s
    pos[:, 2] = lo[2] + iz.ravel().astype(np.float32) * s
    return pos


@dataclasses.dataclass
class PBFParams:
    """Mirror of PBFSettings (include/novaphy/fluid/pbf_solver.h:13-38)."""
    rest_density: float = 1000.0
    kernel_radius: float = 0.1
    particle_radius: float = 0.025
    solver_iterations: int = 4
    epsilon: float = 100.0
    xsph_viscosity: float = 0.01
    vorticity_epsilon: float = 0.0
    corr_delta_q: float = 0.3
    corr_k: float = 0.001
    corr_n: int = 4
    use_domain_bounds: bool = False
    domain_lower: Tuple[float, float, float] = (-10.0, 0.0, -10.0)
    domain_upper: Tuple[float, float, float] = (10.0, 10.0, 10.0)

    def particle_mass(self, spacing: float) -> float:
        rd = np.float32(self.rest_density)
        s = np.float32(spacing)
        return float(rd * s * s * s)


@dataclasses.dataclass
class Scene:
    name: str
    positions: np.ndarray        # (n, 3) float32, original particle order
    velocities: np.ndarray       # (n, 3) float32
    params: PBFParams
    dt: float
    gravity: Tuple[float, float, float]
    mass: float
    steps: int
    seed: int

    @property
    def n(self) -> int:
        return int(self.positions.shape[0])


def _params_dam_break(domain_upper, vorticity=0.0) -> PBFParams:
    return PBFParams(rest_density=1000.0, kernel_radius=0.1, particle_radius=0.025,
                     solver_iterations=4, epsilon=100.0, xsph_viscosity=0.01,
                     vorticity_epsilon=vorticity, corr_delta_q=0.2, corr_k=0.1, corr_n=4,
                     use_domain_bounds=True, domain_lower=(0.0, 0.0, 0.0),
                     domain_upper=tuple(float(x) for x in domain_upper))


def dam_break(nx: int, ny: int, nz: int, domain_upper, seed: int, steps: int,
              vorticity: float = 0.0, name: str = "dam_break",
              lower: Optional[Tuple[float, float, float]] = None) -> Scene:
    """Block of nx*ny*nz particles at spacing 2r with its lower corner at
    (r, r, r) (or ``lower``), jittered by +-0.001 m from ``seed``."""
    r = 0.025
    s = 2 * r
    lo = np.array(lower if lower is not None else (r, r, r), np.float32)
    # upper sits half a spacing past the last lattice point, so the float32
    # floor in the count rule lands on n-1 without rounding doubt.
    hi = lo + np.array([(nx - 0.5) * s, (ny - 0.5) * s, (nz - 0.5) * s], np.float32)
    pos = generate_fluid_block(lo, hi, s)
    assert pos.shape[0] == nx * ny * nz, (pos.shape, nx, ny, nz)
    pos = (pos + jitter(pos.shape[0], seed)).astype(np.float32)
    prm = _params_dam_break(domain_upper, vorticity)
    return Scene(name=name, positions=pos, velocities=np.zeros_like(pos), params=prm,
                 dt=float(np.float32(1.0 / 120.0)), gravity=(0.0, -9.81, 0.0),
                 mass=prm.particle_mass(s), steps=steps, seed=seed)


def dam_break_8k() -> Scene:
    return dam_break(20, 20, 20, (2.0, 1.0, 1.0), seed=0, steps=200, name="dam_break_8k")


def dam_break_1m() -> Scene:
    return dam_break(100, 100, 100, (10.0, 5.0, 5.0), seed=1, steps=500, vorticity=1e-4,
                     name="dam_break_1m")


def coupled_500k_fluid() -> Scene:
    return dam_break(100, 50, 100, (10.0, 5.0, 5.0), seed=2, steps=300, name="coupled_500k")


def batched_scene_sizes(count: int = 1024, seed: int = 3, lo: int = 2000, hi: int = 64000):
    """Per-scene block dims for the batched workload: N log-uniform in [lo, hi],
    aspect ratios in [1, 4], box sized for a 25-50 % initial fill."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        n_target = math.exp(rng.uniform(math.log(lo), math.log(hi)))
        ax = rng.uniform(1.0, 4.0)   # x : z aspect
        ay = rng.uniform(1.0, 4.0)   # y : z aspect
        nz = max(4, int(round((n_target / (ax * ay)) ** (1.0 / 3.0))))
        nx = max(4, int(round(ax * nz)))
        ny = max(4, int(round(ay * nz)))
        fill = rng.uniform(0.25, 0.50)
        if lo <= nx * ny * nz <= hi:  # lattice rounding can step outside the range
            out.append((nx, ny, nz, fill))
    return out


def with_s_corr(sc: Scene, s_corr: str) -> Scene:
    """s_corr "configured": the dam breaks' k = 0.1, n = 4, dq = 0.2h (the
    parameters every workload shares); "default": PBFSettings' defaults
    k = 0.001, dq = 0.3h (pbf_solver.h:21-23), under which the block stays a
    liquid instead of collapsing into stacks."""
    if s_corr == "default":
        sc.params = dataclasses.replace(sc.params, corr_k=0.001, corr_delta_q=0.3)
    elif s_corr != "configured":
        raise ValueError(f"s_corr must be 'configured' or 'default', not {s_corr!r}")
    return sc


def batched_scene(nx: int, ny: int, nz: int, fill: float, seed: int, steps: int = 200,
                  name: str = "batched", s_corr: str = "configured") -> Scene:
    """One scene of the batched / served workloads: a corner block whose box
    keeps the block's footprint in z, doubles height headroom and stretches x
    so that the block volume is ``fill`` of the box.  Same parameters as the
    other dam breaks (r, h, FP32, 4 iterations, XSPH, s_corr)."""
    s = 0.05
    bx, by, bz = nx * s, ny * s, nz * s
    box_z = bz + 0.05
    box_y = by * 1.5 + 0.05
    box_x = max(bx + 0.05, (bx * by * bz) / (fill * box_y * box_z))
    sc = dam_break(nx, ny, nz, (box_x, box_y, box_z), seed=seed, steps=steps, name=name)
    return with_s_corr(sc, s_corr)


def batched_1024(count: int = 1024, steps: int = 200, s_corr: str = "configured") -> List[Scene]:
    dims = batched_scene_sizes(count)
    return [batched_scene(nx, ny, nz, f, seed=3 * 100000 + k, steps=steps,
                          name=f"batched_{k}", s_corr=s_corr) for k, (nx, ny, nz, f) in enumerate(dims)]


def lattice_rest_density(spacing: float = 0.05, h: float = 0.1, rest_density: float = 1000.0):
    """Closed-form interior density of a perfect cubic lattice under poly6 in
    float64: sum over lattice offsets with |r| < h (strict), including self."""
    m = rest_density * spacing ** 3
    k = int(math.ceil(h / spacing)) + 1
    coef = 315.0 / (64.0 * math.pi * h ** 9)
    tot = 0.0
    for i in range(-k, k + 1):
        for j in range(-k, k + 1):
            for l in range(-k, k + 1):
                r2 = (i * i + j * j + l * l) * spacing * spacing
                if r2 < h * h:
                    tot += m * coef * (h * h - r2) ** 3
    return tot


def block_dims(rng: np.random.Generator, lo: int, hi: int):
    """Block dims (nx, ny, nz) with N = nx*ny*nz log-uniform in [lo, hi],
    aspect ratios in [1, 4], and a fill fraction in [0.25, 0.5]."""
    while True:
        n_target = math.exp(rng.uniform(math.log(lo), math.log(hi)))
        ax = rng.uniform(1.0, 4.0)
        ay = rng.uniform(1.0, 4.0)
        nz = max(4, int(round((n_target / (ax * ay)) ** (1.0 / 3.0))))
        nx = max(4, int(round(ax * nz)))
        ny = max(4, int(round(ay * nz)))
        fill = rng.uniform(0.25, 0.50)
        if lo <= nx * ny * nz <= hi:
            return nx, ny, nz, fill


@dataclasses.dataclass
class Sphere:
    radius: float
    density: float
    position: Tuple[float, float, float]


def served_request(k: int, rng: np.random.Generator, lo: int = 4000, hi: int = 256000,
                   max_spheres: int = 8, s_corr: str = "configured") -> Scene:
    """One served request: a dam break (as in the batched workload, with a box
    tall enough for the drop) and 0..max_spheres rigid spheres above it.
    The scene carries ``spheres`` (list of Sphere); the box floor is a ground
    plane for the rigid bodies."""
    nx, ny, nz, fill = block_dims(rng, lo, hi)
    s = 0.05
    bx, by, bz = nx * s, ny * s, nz * s
    box_z = bz + 0.05
    box_y = max(by * 1.5 + 0.05, by + 1.5)
    box_x = max(bx + 0.05, (bx * by * bz) / (fill * box_y * box_z))
    sc = with_s_corr(dam_break(nx, ny, nz, (box_x, box_y, box_z), seed=4 * 1000003 + k, steps=0,
                               name=f"served_{k}"), s_corr)
    spheres = []
    for _ in range(int(rng.integers(0, max_spheres + 1))):
        r = float(rng.uniform(0.1, 0.2))
        x = float(rng.uniform(min(r + 0.05, bx / 2), max(bx - r, bx / 2)))
        z = float(rng.uniform(min(r + 0.05, bz / 2), max(bz - r, bz / 2)))
        y = float(by + rng.uniform(0.3, 1.0))
        spheres.append(Sphere(radius=r, density=float(rng.uniform(300.0, 2000.0)), position=(x, y, z)))
    sc.spheres = spheres
    return sc


def coupled_500k_world(seed: int = 2, flags: int = 0):
    """The rigid-fluid coupling workload as a FluidWorld: the 500k-particle
    block (100x50x100) in the 10x5x5 m box with the dam-break parameters, a
    ground plane sampled with boundary particles over the box, and 64 rigid
    bodies: 32 spheres (r ~ U(0.1, 0.2)) and 32 boxes (half extents
    ~ U(0.05, 0.15)^3), density ~ U(300, 2000), seeded positions over the
    block footprint 0.3-1.0 m above its surface, identity orientation.
    Returns (world, fluid scene)."""
    import novaphy_amd as nv
    sc = coupled_500k_fluid()
    rng = np.random.default_rng(seed)
    b = nv.ModelBuilder()
    b.add_ground_plane(y=0.0)
    top = float(sc.positions[:, 1].max())
    lo = sc.positions.min(0)
    hi = sc.positions.max(0)
    for k in range(64):
        x = rng.uniform(lo[0] + 0.3, hi[0] - 0.3)
        z = rng.uniform(lo[2] + 0.3, hi[2] - 0.3)
        y = top + rng.uniform(0.3, 1.0)
        dens = rng.uniform(300.0, 2000.0)
        t = nv.Transform.from_translation(np.array([x, y, z], np.float32))
        if k < 32:
            r = rng.uniform(0.1, 0.2)
            i = b.add_body(nv.RigidBody.from_sphere(float(dens * 4.0 / 3.0 * np.pi * r ** 3), float(r)), t)
            b.add_shape(nv.CollisionShape.make_sphere(float(r), i))
        else:
            he = rng.uniform(0.05, 0.15, 3).astype(np.float32)
            i = b.add_body(nv.RigidBody.from_box(float(dens * 8.0 * np.prod(he)), he), t)
            b.add_shape(nv.CollisionShape.make_box(he, i))
    world = nv.FluidWorld(b.build(), [empty_block()], nv.SolverSettings(), pbf_settings(sc.params),
                          boundary_extent=10.0, flags=flags)
    world.fluid_state.init_arrays(sc.positions, sc.velocities)
    return world, sc


def pbf_settings(p: PBFParams):
    """novaphy_amd.PBFSettings with the fields of a PBFParams."""
    import novaphy_amd as nv
    prm = nv.PBFSettings()
    for f in ("rest_density", "kernel_radius", "particle_radius", "solver_iterations", "epsilon",
              "xsph_viscosity", "vorticity_epsilon", "corr_delta_q", "corr_k", "corr_n", "use_domain_bounds"):
        setattr(prm, f, getattr(p, f))
    prm.domain_lower = np.array(p.domain_lower, np.float32)
    prm.domain_upper = np.array(p.domain_upper, np.float32)
    return prm


def empty_block():
    """A FluidBlockDef of spacing 0.05 whose particles are replaced with
    ParticleState.init_arrays (it fixes the particle spacing and mass)."""
    import novaphy_amd as nv
    blk = nv.FluidBlockDef()
    blk.particle_spacing = 0.05
    blk.lower = np.zeros(3, np.float32)
    blk.upper = np.zeros(3, np.float32)
    return blk


def window_cells(sc: Scene, margin: int = 4) -> int:
    """Cells of a bounded scene's cell window (domain cells +- margin per
    axis), the cell capacity a batch context needs for it."""
    h = sc.params.kernel_radius
    inv = np.float32(1.0) / np.float32(h)
    lo = [int(np.floor(np.float32(v) * inv)) - margin for v in sc.params.domain_lower]
    hi = [int(np.floor(np.float32(v) * inv)) + margin for v in sc.params.domain_upper]
    return int(np.prod([b - a + 1 for a, b in zip(lo, hi)]))
