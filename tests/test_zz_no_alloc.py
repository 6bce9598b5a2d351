"""pic_moments_noalloc (moments without allocation, partial sums in the
library's static device array) against pic_moments through the C ABI.

Two calls in a row share the GPU's scratch array: the first on a config-0
context (32x8x8 cells, 98,304 particles), the second on a smaller one
(16x4x4 cells).  With the same number of partial blocks as pic_moments
(1024) each result must equal pic_moments' bit for bit; with fewer blocks
(a different summation tree over the same per-particle terms) within 1e-12
of the largest moment of its row.  Both forms; plus the refusals.
"""
import ctypes as C
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

pic = importlib.import_module("0buf_amd")


def _context(precision, scale_n, steps):
    params, parts = pic.generate_workload(0, scale_n=scale_n)
    params.precision = precision
    g = pic.Pic(params)
    for (sort, x, y, z, px, py, pz, m, qm) in parts:
        g.load_particles(sort, x, y, z, px, py, pz, m, qm)
    for nt in range(1, steps + 1):
        g.step(nt)
    return g, sum(len(p[1]) for p in parts)


def _noalloc(g, nblocks):
    import torch
    out = torch.full((27,), float("nan"), dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    rc = g.lib.pic_moments_noalloc(g._ctx, nblocks, C.c_void_p(out.data_ptr()))
    assert rc == 0, g.lib.pic_last_hip_error().decode()
    assert g.lib.pic_sync(g._ctx) == 0
    return out.cpu().numpy().reshape(3, 9)


@pytest.mark.parametrize("precision", [pic.FP64, pic.FP32])
def test_noalloc_moments_match_pic_moments(precision):
    import torch
    assert torch.cuda.is_available()
    big, nbig = _context(precision, None, 3)
    small, nsmall = _context(precision, (16, 4, 4), 2)
    assert nsmall < nbig
    try:
        for g in (big, small):           # two calls in a row, the second smaller
            got = _noalloc(g, 1024)
            want = g.moments()
            assert np.array_equal(got, want), (got - want)
        assert got[:, 0].sum() == nsmall
        # fewer partial blocks: another summation tree, the same sums
        got = _noalloc(small, 7)
        want = small.moments()
        scale = np.abs(want).max(axis=1, keepdims=True)
        assert np.all(np.abs(got - want) <= 1e-12 * scale), (got - want) / scale
        # the larger context again after the small one: no stale partials
        assert np.array_equal(_noalloc(big, 1024), big.moments())
    finally:
        big.close()
        small.close()


def test_noalloc_moments_refusals():
    import torch
    g, _ = _context(pic.FP64, (16, 4, 4), 1)
    out = torch.zeros(27, dtype=torch.float64, device="cuda:0")
    ptr = C.c_void_p(out.data_ptr())
    try:
        for nb in (0, -1, 2049):
            assert g.lib.pic_moments_noalloc(g._ctx, nb, ptr) == 1, nb
        assert g.lib.pic_moments_noalloc(g._ctx, 16, None) == 1
        assert g.lib.pic_moments_noalloc(g._ctx, 2048, ptr) == 0
        assert g.lib.pic_sync(g._ctx) == 0
    finally:
        g.close()
    # particles loaded but not committed: committing would allocate
    params, parts = pic.generate_workload(0, scale_n=(8, 4, 4))
    h = pic.Pic(params)
    try:
        sort, x, y, z, px, py, pz, m, qm = parts[0]
        h.load_particles(sort, x, y, z, px, py, pz, m, qm)
        assert h.lib.pic_moments_noalloc(h._ctx, 16, ptr) == 7
    finally:
        h.close()
