"""The C-ABI library without a GPU: it loads, exports every symbol the header
declares, and refuses compute cleanly (no crash, no CPU fallback)."""
import os
import re

import numpy as np
import pytest

from novaphy_amd import capi, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "novaphy_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(npx_[a-z0-9_]+)\s*\(", txt)))


def test_header_symbols_exported():
    L = capi.load()
    names = _declared()
    assert len(names) >= 27
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(capi.SYMBOLS) == names


def test_create_without_device_fails_cleanly():
    if capi.device_count() > 0:
        pytest.skip("a GPU is present; the no-device path is not reachable")
    with pytest.raises(capi.NpxError) as e:
        capi.Context(scenes.dam_break_8k().params, capacity=16)
    assert e.value.code == capi.NPX_EDEVICE
    assert "device" in str(e.value)


def test_invalid_params_rejected_before_device():
    bad = scenes.PBFParams(kernel_radius=0.0)
    with pytest.raises((ValueError, capi.NpxError)):
        capi.Context(bad, capacity=1)


def test_scene_bytes_admission_estimate():
    sc = scenes.dam_break_8k()
    b = capi.scene_bytes(sc.params, sc.n)
    # 8000 particles -> 8192 slots * (235 B of state and sort scratch + 64-entry
    # skin and in-range lists), plus the cell window (cells -4..24, -4..14, -4..14
    # for the 2x1x1 m domain at h = 0.1) * 16 B
    assert b == 8192 * (235 + 512) + 29 * 19 * 19 * 16
    assert capi.scene_bytes(sc.params, 2 * sc.n) > b


def _shape_model():
    """Boxes and spheres with rotated local transforms on rotated bodies, and
    a tilted plane."""
    import novaphy_amd as nv
    b = nv.ModelBuilder()
    n = np.array([0.1, 1.0, 0.2], np.float32)
    b.add_shape(nv.CollisionShape.make_plane(n / np.linalg.norm(n), 0.05))
    rot = nv.Transform(np.array([0.05, -0.02, 0.01], np.float32), np.array([0.1, 0.2, -0.3, 0.93], np.float32))
    i = b.add_body(nv.RigidBody.from_box(2.0, np.array([0.1, 0.07, 0.13], np.float32)),
                   nv.Transform(np.array([0.4, 0.6, 0.3], np.float32), np.array([0.0, 0.38, 0.0, 0.92], np.float32)))
    b.add_shape(nv.CollisionShape.make_box(np.array([0.1, 0.07, 0.13], np.float32), i, rot))
    j = b.add_body(nv.RigidBody.from_sphere(1.0, 0.17), nv.Transform.from_translation(np.array([0.1, 0.9, 0.2], np.float32)))
    b.add_shape(nv.CollisionShape.make_sphere(0.17, j, rot))
    b.add_shape(nv.CollisionShape.make_sphere(0.11, j))
    return b.build()


def test_boundary_sample_counts_match_host_sampler():
    """npx_sample_boundaries sizes its output on the host with the samplers'
    own expressions: the count equals sample_model_boundaries' (no device
    needed for the size query)."""
    import ctypes as C
    import novaphy_amd as nv
    m = _shape_model()
    shapes, n = capi.shapes_from_model(m)
    assert [shapes[k].type for k in range(n)] == [capi.SHAPE_PLANE, capi.SHAPE_BOX, capi.SHAPE_SPHERE,
                                                  capi.SHAPE_SPHERE]
    assert [shapes[k].body for k in range(n)] == [-1, 0, 1, 1]
    cnt = C.c_int64()
    assert capi.load().npx_sample_boundaries(shapes, n, C.c_float(0.05), C.c_float(0.7), None, None, 0,
                                             C.byref(cnt)) == 0
    assert cnt.value == len(nv.sample_model_boundaries(m, 0.05, 0.7))
    bad = capi.NpxShape(type=7)
    assert capi.load().npx_sample_boundaries(C.byref(bad), 1, C.c_float(0.05), C.c_float(0.7), None, None, 0,
                                             C.byref(cnt)) == capi.NPX_EINVAL


_ENOMEM_CHILD = r"""
import ctypes as C, resource, sys
sys.path.insert(0, sys.argv[1])
from novaphy_amd import capi
L = capi.load()
vm = 0
for line in open('/proc/self/status'):
    if line.startswith('VmSize:'):
        vm = int(line.split()[1]) * 1024
# room for the interpreter to keep running, far too little for the request
resource.setrlimit(resource.RLIMIT_AS, (vm + (1 << 30), resource.RLIM_INFINITY))
shape = capi.NpxShape(type=capi.SHAPE_SPHERE, body=0, radius=0.1)
cnt = C.c_int64()
# 2^27 shapes: the host table of per-shape constants needs ~20 GB
r = L.npx_sample_boundaries(C.byref(shape), 1 << 27, C.c_float(0.05), C.c_float(1.0), None, None, 0, C.byref(cnt))
print(r)
print(L.npx_last_error(None).decode())
"""


def test_host_allocation_failure_returns_enomem():
    """A host allocation that fails inside an entry point comes back as
    NPX_ENOMEM with a message; no C++ exception crosses the C ABI (without
    the guard the child process would abort in std::terminate)."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, "-c", _ENOMEM_CHILD, ROOT], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = r.stdout.strip().splitlines()
    assert int(lines[0]) == capi.NPX_ENOMEM
    assert "npx_sample_boundaries" in lines[1] and "allocation" in lines[1]


def test_capacity_beyond_slot_limit_rejected():
    """Capacities whose slot arithmetic would overflow int32 are refused with
    NPX_EINVAL before any device work; the batch variant checks them at
    creation too."""
    import ctypes as C
    L = capi.load()
    h = C.c_void_p()
    assert L.npx_batch_create(0x7FFFFF00, 1024, 4, 0, C.byref(h)) == capi.NPX_EINVAL
    assert "slots" in L.npx_last_error(None).decode()
    prm = capi.to_params(scenes.dam_break_8k().params)
    assert L.npx_create(C.byref(prm), 0x7FFFFFF0, 0, C.byref(h)) == capi.NPX_EINVAL
    assert not h.value
