"""npx_grid_query_noalloc (the allocation-free variant of npx_grid_query)
against npx_grid_query through the C ABI: candidate counts and lists equal
bit for bit, on two calls in a row, the second with a smaller input, and the
size limits refused with a clear error."""
import ctypes as C

import numpy as np
import pytest

from novaphy_amd import capi

pytestmark = pytest.mark.gpu


def _noalloc(points, cell, queries, cap):
    L = capi.load()
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    q = np.ascontiguousarray(queries, np.float32).reshape(-1, 3)
    counts = np.full(q.shape[0], -7, np.int32)
    out = np.full(max(cap, 1), -7, np.int32)
    tot = C.c_int64(-1)
    r = L.npx_grid_query_noalloc(pts.ctypes.data_as(C.c_void_p), pts.shape[0], C.c_float(cell),
                                 q.ctypes.data_as(C.c_void_p), q.shape[0], counts.ctypes.data_as(C.c_void_p),
                                 out.ctypes.data_as(C.c_void_p), cap, C.byref(tot))
    return r, counts, out, tot.value


def _case(seed, n, nq, span):
    rng = np.random.default_rng(seed)
    pts = rng.uniform(-span, span, (n, 3)).astype(np.float32)
    # queries near the origin too, where Teschner keys alias and buckets repeat
    q = np.concatenate([rng.uniform(-span, span, (nq - 8, 3)),
                        rng.uniform(-0.15, 0.15, (8, 3))]).astype(np.float32)
    return pts, q


def test_noalloc_grid_query_matches_allocating_entry_point(gpu):
    cell = 0.1
    for seed, n, nq, span in ((11, 20000, 4000, 1.2), (12, 3000, 700, 0.6)):  # second call smaller
        print(f"case: grid_query_noalloc seed={seed} n={n} nq={nq}", flush=True)
        pts, q = _case(seed, n, nq, span)
        ref_counts, ref_lists = capi.grid_query(pts, cell, q)
        r, counts, out, tot = _noalloc(pts, cell, q, int(ref_lists.shape[0]) + 10)
        assert r == capi.NPX_OK, capi.load().npx_last_error(None).decode()
        np.testing.assert_array_equal(counts, ref_counts)
        assert tot == ref_lists.shape[0]
        np.testing.assert_array_equal(out[:tot], ref_lists)
        assert np.all(out[tot:] == -7)  # nothing written past the total
        # counts only (no lists requested): same counts and total
        r, counts2, _, tot2 = _noalloc(pts, cell, q, 0)
        assert r == capi.NPX_OK and tot2 == tot
        np.testing.assert_array_equal(counts2, ref_counts)


def test_noalloc_grid_query_limits(gpu):
    L = capi.load()
    pts = np.zeros((70000, 3), np.float32)
    r, _, _, _ = _noalloc(pts, 0.1, pts[:4], 0)
    assert r == capi.NPX_EINVAL and "65536 points" in L.npx_last_error(None).decode()
    far = np.array([[0, 0, 0], [100.0, 100.0, 100.0]], np.float32)
    r, _, _, _ = _noalloc(far, 0.1, far[:1], 0)
    assert r == capi.NPX_EDOMAIN and "2^18 cells" in L.npx_last_error(None).decode()
    # still usable after a refusal
    pts, q = _case(13, 500, 100, 0.3)
    ref_counts, ref_lists = capi.grid_query(pts, 0.1, q)
    r, counts, out, tot = _noalloc(pts, 0.1, q, int(ref_lists.shape[0]))
    assert r == capi.NPX_OK
    np.testing.assert_array_equal(counts, ref_counts)
    np.testing.assert_array_equal(out[:tot], ref_lists)
