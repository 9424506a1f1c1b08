#!/usr/bin/env python3
"""Benchmark of the MI355X PBF step (one JSON line on stdout).

Default workload: the single large dam break -- 100x100x100 = 1M FP32
particles, r = 0.025 m, h = 0.1 m, rest density 1000, 10x5x5 m box (the
100-particle lattice spans 4.95 m, so the box keeps the 2:1:1 aspect at
10x5x5 m), seeded +-1 mm jitter, dt = 1/120 s, 4 iterations, XSPH c = 0.01,
vorticity confinement eps = 1e-4, s_corr k = 0.1, n = 4, dq = 0.2h (the
headline `value`).  At k = 0.1 the tensile correction overshoots and the
fluid collapses into clusters of coincident particles, so the step cost
grows with the cluster size; the same scene with PBFSettings' default s_corr
(k = 0.001, dq = 0.3h) stays a liquid and is reported beside it under
"default_s_corr" (--s-corr picks one or both).

A step is one PBFSolver::step over all particles (predict, grid, 4 x
(density+lambda, delta-p), velocity update, vorticity, XSPH, commit).  The
timed region is the first K steps (default 500, the scene's specified run)
from the initial state, after W untimed warm-up steps on the same context;
inputs are resident on the device, the region is bracketed by device
synchronisation, and each step is one hipGraph replay.  Per-kernel
durations come from HIP events recorded on the context's stream around
every launch of a second run of the same K steps from the same initial
state (events need stream launches; window_ms_per_step gives the step time
across that run, and its final state is checked to be bitwise the timed
run's).  The default run also reports, as extra keys, the
batched workload (1024 scenes; the batch and the plain per-scene loop timed
on the same step window) and two served samples from the seeded Poisson
stream: the requests of the first 15 s at 4/s (light load: p50/p99 step
latency; the rate follows the arrivals) and 300 requests at 30/s, which
arrive faster than they are served (the rate is the server's capacity).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload 1m|8k|batch|coupled|serve]
                  [--s-corr both|configured|default] [--extras 0|1] [--no-graph] [--copy-calibration]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "PBF particle-steps/s at 1M + % HBM roofline; batched p50/p99 step latency"
HBM_PEAK_GBPS = 8000.0
# Algorithmic (compulsory) FP32/int32 bytes per particle for each pass, 12 B
# per 3-vector, each dependency-forced pass reading and writing each array
# once: predict + cell key 40, stable counting sort 13, reorder 56; per
# iteration density + lambda 20 (x* in; rho, lambda out) and delta-p + apply
# 28 (x*, lambda in; x* out); velocity + XSPH + clamp + commit 52 (x*, x, rho
# in; x, v out), plus 48 with vorticity (omega pass: x*, v in, omega out 36;
# the final pass reads omega, +12).  The fused omega_xsph / vort_commit
# kernels split the last two: omega_xsph 36 with vorticity (24 without: the
# x*, v reads of the XSPH sum) and vort_commit 64 (28 without), so a step is
# 161 + 48 I bytes (+48 with vorticity).
def pass_bytes(vort: bool):
    return {"density_lambda": 20, "density_lambda_build": 20, "delta": 28,
            "omega_xsph": 36 if vort else 24, "vort_commit": 64 if vort else 28,
            "predict_bin": 40, "gather": 56}


def fold_dense(kt):
    """Kernel times with the build pass's other launches added to its
    list-free per-lane kernel ("density_lambda_build"): the dense-particle
    kernel (name + "_dense": the crowded particles, one wave per batch) and
    the list-building variant (name + "_lists", the scenes that build skin
    lists this step; each block of either per-lane launch leaves at once when
    its scene is the other's).  Together they are one pass over every
    particle.  The later passes sweep their crowded particles in the first
    blocks of their own launch.  Launch counts stay those of the per-lane
    kernel."""
    sub = ("_dense", "_lists")
    out = {k: list(v) for k, v in kt.items() if not any(k.endswith(x) and k[:-len(x)] in kt for x in sub)}
    for k, v in kt.items():
        for x in sub:
            if k.endswith(x) and k[:-len(x)] in kt:
                out[k[:-len(x)]][0] += v[0]
    return {k: tuple(v) for k, v in out.items()}


def step_bytes(iters: int, vort: bool) -> int:
    return 161 + 48 * iters + (48 if vort else 0)


def profiled_traffic(variant: str, kernel: str):
    """HBM bytes per launch of `kernel` in the 1M workload `variant`, from the
    calibrated rocprofv3 FETCH_SIZE / WRITE_SIZE runs summarised in
    profiles/hbm_traffic_1m.md (None when not profiled)."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_1m.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 4 and f[0] == "traffic" and f[1] == variant and f[2] == kernel:
            return int(f[3])
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profiled_counters(variant: str, kernel: str):
    """VALU busy fraction and active lanes per VALU instruction of `kernel` in
    the 1M workload `variant`, from the rocprofv3 SQ counter runs summarised in
    profiles/pbf_1m_valu_counters.md (None when not profiled)."""
    path = os.path.join(ROOT, "profiles", "pbf_1m_valu_counters.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 5 and f[0] == "valu" and f[1] == variant and f[2] == kernel:
            return {"valu_busy": float(f[3]), "active_lanes_per_valu_instr": float(f[4])}
    return None


def window_ms(trace, steps, window):
    """Mean ms per step over consecutive windows of `window` steps, from the
    kernel trace of the timed region (each step starts with its first
    kernel, "fill_counts")."""
    starts = [st for name, st, _ in trace if name == "fill_counts"]
    last_end = max(st + du for _, st, du in trace)
    starts = starts[:steps] + [last_end]
    out = []
    for a in range(0, len(starts) - 1, window):
        b = min(a + window, len(starts) - 1)
        out.append({"steps": f"{a + 1}-{b}", "ms_per_step": round((starts[b] - starts[a]) / (b - a), 3)})
    return out


def run_single(sc, steps, warmup, flags, cpu_seconds, variant=None, kernel_timing=True):
    """The timed region is the first `steps` steps of the workload from its
    initial state (after `warmup` untimed steps on the same context, which is
    then reset to the initial state), each step one hipGraph replay as in
    normal use.  The per-kernel durations of the roofline come from a second
    run of the same steps from the same initial state, with a HIP event pair
    recorded on the context's stream around every launch: events need stream
    launches, which cost about a tenth more than graph replay, so that run is
    not the timed one.  The steps are deterministic; the final states of the
    two runs are compared bit for bit (`kernel_run_identical`)."""
    from novaphy_amd import capi
    n = sc.n
    ctx = capi.Context(sc.params, capacity=n, flags=flags)
    ctx.set_particles(sc.positions, sc.velocities)
    log(f"[bench] {sc.name}: n={n}, warmup {warmup} steps")
    ctx.step(sc.dt, sc.gravity, sc.mass, warmup)
    ctx.synchronize()
    ctx.set_particles(sc.positions, sc.velocities)  # back to the initial state
    start_state = (sc.positions, sc.velocities)
    ctx.synchronize()
    log(f"[bench] timed region: steps 1-{steps} from the initial state")
    t0 = time.perf_counter()
    ctx.step(sc.dt, sc.gravity, sc.mass, steps)
    ctx.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    value = n * steps / wall
    if not kernel_timing:  # counter-collection runs: no events around the launches
        ctx.close()
        return {"value": value, "ms_per_step": wall * 1e3 / steps,
                "window_ms_per_step": None, "roofline": None, "step_roofline": None, "kernels": None}
    final = ctx.get_particles()
    log(f"[bench] kernel-timed run: steps 1-{steps} again from the initial state, HIP events around every launch")
    ctx.set_particles(sc.positions, sc.velocities)
    ctx.set_kernel_timing(True)
    ctx.synchronize()
    t2 = time.perf_counter()
    ctx.step(sc.dt, sc.gravity, sc.mass, steps)
    ctx.synchronize()
    t3 = time.perf_counter()
    final2 = ctx.get_particles()
    identical = all(np.array_equal(a, b) for a, b in zip(final, final2))
    kt_raw = ctx.kernel_times()
    trace = ctx.kernel_trace()
    kt = fold_dense(kt_raw)
    ctx.set_kernel_timing(False)
    iters = sc.params.solver_iterations
    vort = sc.params.vorticity_epsilon != 0
    PASS_BYTES = pass_bytes(vort)
    main = max((k for k in kt if k in PASS_BYTES), key=lambda k: kt[k][0])
    ms_tot, launches = kt[main]
    avg_s = ms_tot / 1e3 / launches
    achieved = PASS_BYTES[main] * n / avg_s / 1e9
    B = step_bytes(iters, vort)
    full = variant and n == 1000000 and not flags
    out = {
        "value": value,
        "ms_per_step": wall * 1e3 / steps,
        "window_ms_per_step": window_ms(trace, steps, max(1, steps // 10)),
        "kernel_run_ms_per_step": round((t3 - t2) * 1e3 / steps, 4),
        "kernel_run_identical": identical,
        "roofline": {"bound": "hbm", "kernel": main,
                     "kernels": [k for k in (main, main + "_lists", main + "_dense") if k in kt_raw],
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 5),
                     "traffic": profiled_traffic(variant, main) if full else None,
                     "bytes_per_particle": PASS_BYTES[main], "particles_per_launch": n,
                     "avg_launch_us": round(avg_s * 1e6, 2),
                     # no MFMA work: the HBM roof is the kernel's roofline; what limits it
                     # is VALU issue and gather latency (SQ counters, profiles/)
                     "limiter": "valu issue / gather latency",
                     "counters": profiled_counters(variant, main) if full else None},
        "step_roofline": {"bytes_per_particle_step": B, "achieved": round(B * value / 1e9, 2),
                          "unit": "GB/s", "frac": round(B * value / 1e9 / HBM_PEAK_GBPS, 5)},
        "kernels": {k: {"total_ms": round(v[0], 3), "launches": v[1],
                        "avg_us": round(v[0] * 1e3 / max(v[1], 1), 2)} for k, v in kt_raw.items()},
    }
    ctx.close()
    if cpu_seconds > 0:
        out["cpu_baseline"] = cpu_baseline(sc, start_state, cpu_seconds)
    return out


def copy_calibration(nbytes=256 << 20, reps=3):
    """Device-to-device copies of a known size through the HIP runtime: the
    reference reading for the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        if hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    wall = time.perf_counter() - t0
    hip.hipFree(src)
    hip.hipFree(dst)
    return {"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "copies": reps,
            "copy_GBps_read_plus_write": round(2 * nbytes * reps / wall / 1e9, 1)}


def run_batch(count, steps, warmup, flags, loop, s_corr="configured"):
    """Batched scenes (one launch per pass over every scene, each step one
    hipGraph replay) against the plain per-scene loop of single-scene
    contexts, both timed on the same step window: steps warmup+1 ..
    warmup+steps from the initial state.  The launch tunables are chosen by
    the autotuner on a separate batch of the same scenes first."""
    from novaphy_amd import autotune, capi, scenes
    scs = scenes.batched_1024(count=count, s_corr=s_corr)
    total = sum(s.n for s in scs)
    slots = sum(((s.n + 255) // 256) * 256 for s in scs)
    cells = sum(scenes.window_cells(s) for s in scs)
    log(f"[bench] batch: {count} scenes, {total} particles, {cells} cells, s_corr {s_corr}")

    def make_batch():
        b = capi.Batch(slot_capacity=slots, cell_capacity=cells, max_scenes=count, flags=flags)
        for s in scs:
            b.add(s.params, s.dt, s.gravity, s.mass, s.positions, s.velocities)
        return b

    b = make_batch()
    best, tried = autotune.autotune(b, steps_per_candidate=1)
    b.close()
    log(f"[bench] autotuner: block {best[0]}, xcd order {best[1]}, skin {best[2]}h, persistent {best[3]} "
        f"({', '.join(f'{k[0]}/{k[1]}/{k[2]}/{k[3]}: {v:.1f} ms' for k, v in tried.items())})")
    b = make_batch()
    b.set_tuning(*best)
    b.step(warmup)
    b.synchronize()
    t0 = time.perf_counter()
    ms = b.step_timed(steps)
    wall = time.perf_counter() - t0
    b.close()
    value = total * steps / wall
    out = {"value": value, "ms_per_step": wall * 1e3 / steps, "particles": total, "scenes": count,
           "window": f"steps {warmup + 1}-{warmup + steps}",
           "p50_step_ms": float(np.percentile(ms, 50)), "p99_step_ms": float(np.percentile(ms, 99)),
           "tuning": {"block": best[0], "xcd_order": best[1], "skin": best[2], "persistent": best[3]}}
    if not loop:
        out.update(loop_value=None, loop_ms_per_step=None, speedup_vs_loop=None)
        return out
    log(f"[bench] per-scene loop: {count} contexts, steps {warmup + 1}-{warmup + steps}")
    ctxs = []
    for s in scs:
        c = capi.Context(s.params, capacity=s.n, flags=flags)
        c.set_tuning(*best)
        c.set_particles(s.positions, s.velocities)
        ctxs.append((c, s))
    for c, s in ctxs:
        c.step(s.dt, s.gravity, s.mass, warmup)
    for c, _ in ctxs:
        c.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        for c, s in ctxs:
            c.step(s.dt, s.gravity, s.mass, 1)
    for c, _ in ctxs:
        c.synchronize()
    lwall = time.perf_counter() - t0
    for c, _ in ctxs:
        c.close()
    out["loop_value"] = total * steps / lwall
    out["loop_ms_per_step"] = lwall * 1e3 / steps
    out["speedup_vs_loop"] = value / out["loop_value"]
    return out


def coupled_world(seed=2, flags=0):
    """The rigid-fluid coupling workload (scenes.coupled_500k_world)."""
    from novaphy_amd import scenes
    return scenes.coupled_500k_world(seed, flags)


def run_coupled(steps, warmup, phase_steps=100, flags=0):
    """FluidWorld.step over the coupled workload: the timed steps run with the
    performance monitor off (the PBF step replays its graph); then
    phase_steps more steps with the monitor on (per-launch HIP events, stream
    launches, a synchronisation after the PBF step) give the phase averages."""
    world, sc = coupled_world(flags=flags)
    log(f"[bench] coupled: {sc.n} fluid particles, {world.num_boundary_particles} boundary particles, 64 bodies")
    dt = np.float32(sc.dt)
    for _ in range(warmup):
        world.step(dt)
    mon = world.performance_monitor
    mon.enabled = False
    t0 = time.perf_counter()
    for _ in range(steps):
        world.step(dt)
    wall = time.perf_counter() - t0
    mon.enabled = True
    mon.reset()
    t1 = time.perf_counter()
    for _ in range(phase_steps):
        world.step(dt)
    mon_ms = (time.perf_counter() - t1) * 1e3 / max(phase_steps, 1)
    phases = {p.name: round(p.avg_ms, 3) for p in mon.phase_stats()}
    return {"value": sc.n * steps / wall, "ms_per_step": wall * 1e3 / steps, "particles": sc.n,
            "boundary_particles": world.num_boundary_particles, "phases_avg_ms": phases,
            "phase_steps": phase_steps, "monitored_ms_per_step": mon_ms}


def run_serve(seconds, rate_hz, threads, backend="fused", n_requests=0, kernel_window=(0, 0), max_seconds=0.0):
    """Served rollouts: Poisson arrivals (seed 4) of dam breaks with N
    log-uniform in [4k, 256k], horizons uniform in [50, 500] and 0-8 rigid
    spheres, admitted against 80 % of 288 GB.  fused: every request shares one
    batch context, with the coupling of all requests with spheres in one
    launch per pass and their boundary particles sampled on the device; mixed:
    requests with spheres run as FluidWorld instances stepped concurrently.
    The sample is the requests arriving in the first `seconds` of server time
    (or the first `n_requests` of the stream); the run ends when they have
    all finished, at 3x their arrival window or after max_seconds (when
    positive), whichever comes first.  kernel_window (a, b): per-kernel
    HIP-event times of the batch over server rounds a..b-1 (none when a >= b)."""
    from novaphy_amd import capi, scenes, serve
    n_req = int(n_requests) if n_requests > 0 else max(1, int(round(seconds * rate_hz)))
    # requests are drawn from the seeded stream as they arrive, so only the
    # scenes that have arrived are held on the host
    stream = serve.poisson_stream(n_req, rate_hz, seed=4, make_scene=lambda k, rng: scenes.served_request(k, rng))
    log(f"[bench] serve: {n_req} requests at {rate_hz}/s")
    b = capi.Batch(slot_capacity=1 << 24, cell_capacity=1 << 26, max_scenes=256)
    be = (serve.FusedBackend(b, 256, device_sampling=True) if backend == "fused"
          else serve.MixedBackend(b, 256, threads=threads))
    srv = serve.Server(be, hbm_budget_bytes=int(0.8 * 288e9), max_scenes=256, bytes_fn=capi.scene_bytes,
                       keep_results=False)
    t0 = time.perf_counter()
    srv.clock = lambda: time.perf_counter() - t0
    nxt = next(stream, None)
    ktimes = None
    sizes, with_spheres, window = [], 0, 0.0
    active = []
    next_log = 30.0
    limit = 3.0 * max(n_req / rate_hz, 1.0)
    if max_seconds > 0:
        limit = min(limit, max_seconds)
    while (nxt is not None or srv.pending or srv.active) and time.perf_counter() - t0 < limit:
        now = srv.clock()
        while nxt is not None and nxt.arrival <= now:
            sizes.append(nxt.scene.n)
            with_spheres += bool(nxt.scene.spheres)
            window = nxt.arrival
            srv.submit(nxt)
            nxt = next(stream, None)
        if srv.steps == kernel_window[0] and kernel_window[0] < kernel_window[1] and ktimes is None:
            b.set_kernel_timing(True)
            ktimes = {}
        if not srv.step():
            time.sleep(0.005)
        if srv.steps == kernel_window[1] and ktimes == {}:
            ktimes = {k: {"total_ms": round(v[0], 3), "launches": v[1]} for k, v in b.kernel_times().items()}
            b.set_kernel_timing(False)
        active.append(len(srv.active))
        el = time.perf_counter() - t0
        if el >= next_log:
            next_log += 30.0
            log(f"[bench] serve t={el:.0f}s: {len(sizes)} arrived, {len(srv.finished)} finished, "
                f"{len(srv.active)} active, {len(srv.pending)} waiting, "
                f"{srv.particle_steps / el / 1e6:.1f} M particle-steps/s so far")
    wall = time.perf_counter() - t0
    if backend != "fused":
        be.close()
    b.close()
    done = [r for r in srv.finished.values() if r.status == serve.DONE]
    lat = srv.latency_percentiles()
    return {"value": srv.particle_steps / wall, "wall_s": wall, "requests": n_req, "done": len(done),
            "unfinished": len(srv.active) + len(srv.pending) + (n_req - len(sizes)),
            "failed": sum(r.status == serve.FAILED for r in srv.finished.values()),
            "p50_step_ms": lat["p50_ms"], "p99_step_ms": lat["p99_ms"], "latency_samples": lat["samples"],
            "mean_active": float(np.mean(active)) if active else 0.0, "rounds": srv.steps,
            "arrival_window_s": window, "admission_s": srv.admit_s, "step_s": srv.busy_s,
            "n_range": [min(sizes), max(sizes)] if sizes else None, "with_spheres": with_spheres,
            "phases_s": {k: round(v, 3) for k, v in getattr(be, "phase_s", {}).items()},
            "kernel_window": (f"rounds {kernel_window[0]}-{kernel_window[1] - 1}" if ktimes else None),
            "kernels": ktimes or None}


def cpu_baseline(sc, state, budget_s):
    """The oracle (serial FP32 restatement of the reference's CPU solver, same
    unordered_map grid and candidate vectors) on one host core, from the GPU
    timed region's starting state: one untimed warm-up step, then as many
    timed steps as fit in the budget (at least one)."""
    from oracle import oracle as O
    O.build()
    pos, vel = state[0], state[1]
    sim = O.Sim(sc.params, pos, vel)
    tw = time.perf_counter()
    sim.step(sc.dt, sc.gravity, sc.mass, 1)  # warm-up
    warm = time.perf_counter() - tw
    t0 = time.perf_counter()
    steps = 0
    while True:
        sim.step(sc.dt, sc.gravity, sc.mass, 1)
        steps += 1
        if time.perf_counter() - t0 >= budget_s or steps >= 10:
            break
    wall = time.perf_counter() - t0
    sim.close()
    try:
        aff = len(os.sched_getaffinity(0))
    except Exception:
        aff = os.cpu_count()
    return {"value": sc.n * steps / wall, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"steps 2-{steps + 1} of the same {sc.n}-particle scene from its initial state, serial, "
                      f"after 1 untimed warm-up step ({warm:.1f} s); {wall:.1f} s timed.  Fewer than 10 timed "
                      f"steps when one step exceeds the {budget_s:.0f} s CPU budget of the run "
                      f"(host: {aff} cores available, 1 used: the reference is serial)"}


def batch_summary(r, s_corr):
    return {"workload": f"batched_{r['scenes']}: {r['scenes']} dam breaks, N log-uniform 2k-64k "
                        f"({r['particles']} particles), s_corr {s_corr}, one launch per pass over all scenes, "
                        "each step one hipGraph replay",
            "value": r["value"], "ms_per_step": r["ms_per_step"], "window": r["window"],
            "p50_step_ms": r["p50_step_ms"], "p99_step_ms": r["p99_step_ms"],
            "per_scene_loop": {"value": r["loop_value"], "ms_per_step": r["loop_ms_per_step"],
                               "window": r["window"]},
            "speedup_vs_per_scene_loop": None if r["speedup_vs_loop"] is None else round(r["speedup_vs_loop"], 2),
            "tuning": r["tuning"]}


def serve_summary(r, rate, backend):
    return {"workload": f"served: the first {r['requests']} of the seeded Poisson request stream at {rate}/s "
                        "(seed 4; N log-uniform 4k-256k, horizon 50-500, 0-8 spheres with Akinci coupling), "
                        f"admission against 80% of 288 GB, {backend} coupling"
                        + (", boundaries sampled on the device" if backend == "fused" else ""),
            "value": r["value"], "p50_step_ms": r["p50_step_ms"], "p99_step_ms": r["p99_step_ms"],
            "serve": {k: r[k] for k in ("requests", "done", "unfinished", "failed", "wall_s", "latency_samples",
                                        "mean_active", "arrival_window_s", "admission_s", "step_s", "n_range",
                                        "with_spheres", "phases_s", "kernel_window", "kernels")}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500,
                    help="timed steps from the initial state (the 1M scene is specified for 500 steps)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="1m", choices=["1m", "8k", "batch", "coupled", "serve"])
    ap.add_argument("--extras", type=int, default=1,
                    help="1m workload: also measure the batched and served workloads (extra keys)")
    ap.add_argument("--serve-seconds", type=float, default=15.0, help="arrival window of the served sample")
    ap.add_argument("--serve-rate", type=float, default=4.0, help="Poisson arrival rate (requests/s)")
    ap.add_argument("--serve-saturated-rate", type=float, default=30.0,
                    help="arrival rate of the default run's saturated served sample (requests/s)")
    ap.add_argument("--serve-saturated-requests", type=int, default=300,
                    help="requests of the default run's saturated served sample")
    ap.add_argument("--serve-kernel-window", default="0,0",
                    help="serve workload: per-kernel times over server rounds a,b (a >= b: off)")
    ap.add_argument("--serve-max-seconds", type=float, default=0.0,
                    help="serve workload: stop serving after this many seconds (0: 3x the arrival window)")
    ap.add_argument("--serve-requests", type=int, default=0,
                    help="serve workload: the first N requests of the stream (0: those of --serve-seconds)")
    ap.add_argument("--serve-threads", type=int, default=8, help="host threads stepping coupled requests (mixed)")
    ap.add_argument("--serve-backend", default="fused", choices=["fused", "mixed"],
                    help="coupling of requests with spheres: fused into the batch, or separate FluidWorlds")
    ap.add_argument("--scenes", type=int, default=1024, help="scene count of the batch workload")
    ap.add_argument("--batch-steps", type=int, default=20,
                    help="timed steps of the batch and of the per-scene loop (the same window)")
    ap.add_argument("--loop", type=int, default=1, help="also time the per-scene loop baseline (0: skip)")
    ap.add_argument("--batch-s-corr", default="configured", choices=["configured", "default"])
    ap.add_argument("--fast-math", action="store_true",
                    help="FMA/rsqrt kernels (within tolerance) instead of the reference operation order")
    ap.add_argument("--no-graph", action="store_true", help="launch kernels directly instead of hipGraph replay")
    ap.add_argument("--kernel-timing", type=int, default=1,
                    help="HIP events around every launch of the timed region (0: none; no roofline then)")
    ap.add_argument("--corr-k", type=float, default=None,
                    help="override the s_corr k of the workload (the configured value is 0.1)")
    ap.add_argument("--cpu-seconds", type=float, default=30.0,
                    help="budget of the host CPU baseline sample of the headline scene (0 disables)")
    ap.add_argument("--s-corr", default="both", choices=["both", "configured", "default"],
                    help="single-scene workloads: the configured s_corr (k=0.1, dq=0.2h; the headline), "
                         "PBFSettings' default (k=0.001, dq=0.3h), or both")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 256 MiB device-to-device copies (reference for the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported: every GPU "
              "box of this build has one GPU (how the path would shard is described in DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        print(json.dumps(copy_calibration()), flush=True)
        return
    from novaphy_amd import capi, scenes
    if capi.device_count() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
    flags = (capi.FLAG_FAST_MATH if args.fast_math else 0) | (capi.FLAG_NO_GRAPH if args.no_graph else 0)
    common = {"metric": METRIC, "unit": "particle-steps/s", "n_gpus": 1, "higher_is_better": True,
              "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic"}
    if args.workload == "batch":
        r = run_batch(args.scenes, args.batch_steps, args.warmup, flags, args.loop, args.batch_s_corr)
        bs = batch_summary(r, args.batch_s_corr)
        line = {**common, "value": r["value"], "steps": args.batch_steps, "warmup": args.warmup,
                "ms_per_step": r["ms_per_step"], "config": {"workload": bs.pop("workload"),
                                                            "particles": r["particles"]}, **bs}
        print(json.dumps(line), flush=True)
        return
    if args.workload == "serve":
        r = run_serve(args.serve_seconds, args.serve_rate, args.serve_threads, args.serve_backend,
                      args.serve_requests, tuple(int(x) for x in args.serve_kernel_window.split(",")),
                      args.serve_max_seconds)
        ss = serve_summary(r, args.serve_rate, args.serve_backend)
        line = {**common, "value": r["value"], "steps": r["rounds"], "warmup": 0, "ms_per_step": r["p50_step_ms"],
                "config": {"workload": ss.pop("workload")}, **ss}
        print(json.dumps(line), flush=True)
        return
    if args.workload == "coupled":
        r = run_coupled(args.steps, args.warmup,
                        flags=capi.FLAG_NO_GRAPH if args.no_graph else 0)
        line = {**common, "value": r["value"], "steps": args.steps, "warmup": args.warmup,
                "ms_per_step": r["ms_per_step"],
                "config": {"workload": f"coupled_500k: {r['particles']} fluid particles + 64 rigid bodies "
                                       f"({r['boundary_particles']} Akinci boundary particles incl. ground "
                                       "plane), FluidWorld.step with host rigid pipeline",
                           "particles": r["particles"]},
                "phases_avg_ms": r["phases_avg_ms"],
                "phases_window": f"steps {args.warmup + args.steps + 1}-{args.warmup + args.steps + r['phase_steps']} "
                                 "with the performance monitor on (per-launch events, no graph replay): "
                                 f"{r['monitored_ms_per_step']:.3f} ms/step there"}
        print(json.dumps(line), flush=True)
        return
    import dataclasses
    base = scenes.dam_break_1m() if args.workload == "1m" else scenes.dam_break_8k()
    if args.corr_k is not None:
        base = dataclasses.replace(base, params=dataclasses.replace(base.params, corr_k=args.corr_k))
    variants = []
    if args.s_corr in ("both", "configured"):
        variants.append(("configured", base))
    if args.s_corr in ("both", "default"):
        dflt = scenes.PBFParams()
        variants.append(("default", dataclasses.replace(base, params=dataclasses.replace(
            base.params, corr_k=dflt.corr_k, corr_delta_q=dflt.corr_delta_q, corr_n=dflt.corr_n))))
    results = [(name, sc, run_single(sc, args.steps, args.warmup, flags, args.cpu_seconds if i == 0 else 0, name,
                                     kernel_timing=bool(args.kernel_timing)))
               for i, (name, sc) in enumerate(variants)]

    def describe(sc):
        p = sc.params
        return (f"{sc.name}: {sc.n} FP32 particles, box {p.domain_upper}, h={p.kernel_radius}, "
                f"{p.solver_iterations} iterations, vorticity {p.vorticity_epsilon}, "
                f"s_corr k={p.corr_k} n={p.corr_n} dq={p.corr_delta_q}h, dt=1/120, steps 1-{args.steps} "
                f"from the initial state, {'fast-math' if args.fast_math else 'reference operation order'}")

    name, sc, r = results[0]
    line = {**common, "value": r["value"], "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": r["ms_per_step"],
            "config": {"workload": describe(sc), "particles": sc.n, "s_corr": name},
            "window_ms_per_step": r["window_ms_per_step"],
            "kernel_run_ms_per_step": r.get("kernel_run_ms_per_step"),
            "kernel_run_identical": r.get("kernel_run_identical"),
            "roofline": r["roofline"], "step_roofline": r["step_roofline"], "kernels": r["kernels"]}
    if "cpu_baseline" in r:
        line["cpu_baseline"] = r["cpu_baseline"]
        line["speedup_vs_cpu"] = round(r["value"] / r["cpu_baseline"]["value"], 1)
    for name2, sc2, r2 in results[1:]:
        line[f"{name2}_s_corr"] = {"workload": describe(sc2), "value": r2["value"], "ms_per_step": r2["ms_per_step"],
                                   "window_ms_per_step": r2["window_ms_per_step"],
                                   "kernel_run_ms_per_step": r2.get("kernel_run_ms_per_step"),
                                   "kernel_run_identical": r2.get("kernel_run_identical"),
                                   "roofline": r2["roofline"],
                                   "step_roofline": r2["step_roofline"], "kernels": r2["kernels"]}
    if args.workload == "1m" and args.extras:
        rb = run_batch(args.scenes, args.batch_steps, args.warmup, flags, args.loop, args.batch_s_corr)
        line["batched"] = batch_summary(rb, args.batch_s_corr)
        rs = run_serve(args.serve_seconds, args.serve_rate, args.serve_threads, args.serve_backend)
        line["served"] = serve_summary(rs, args.serve_rate, args.serve_backend)
        # the light-load sample above is arrival-bound (its rate is the offered load); this one
        # arrives faster than the GPU serves, so its rate is the server's capacity
        rs2 = run_serve(0, args.serve_saturated_rate, args.serve_threads, args.serve_backend,
                        n_requests=args.serve_saturated_requests)
        line["served_saturated"] = serve_summary(rs2, args.serve_saturated_rate, args.serve_backend)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
