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
    # 8000 particles -> 8192 slots * (217 B of state + 64-entry skin and in-range lists),
    # plus the cell window (cells -4..24, -4..14, -4..14 for the 2x1x1 m domain
    # at h = 0.1) * 16 B
    assert b == 8192 * (217 + 512) + 29 * 19 * 19 * 16
    assert capi.scene_bytes(sc.params, 2 * sc.n) > b
