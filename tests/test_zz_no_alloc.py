"""asc_ncc_noalloc (no allocation: static device scratch) against asc_ncc on
the same buffers, through the C ABI: the nine moments and the NCC are the
same bits. Two calls in a row, the second on a smaller input (the scratch
keeps the first call's partials, which the second must not read); sizes
above the stated maximum are refused before anything is launched."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(rng, n):
    f = rng.random(n).astype(np.float32)
    g = (0.3 * f + 0.7 * rng.random(n)).astype(np.float32)
    m = (rng.random(n) > 0.2).astype(np.float32)
    return f * m, g * m, m


def test_noalloc_ncc_matches_asc_ncc(asc, ctx):
    rng = np.random.default_rng(77)
    out = ctx.alloc(10 * 8)
    for n in (1_500_001, 40_961, 3):
        f, g, m = _inputs(rng, n)
        F, G, M = ctx.to_device(f), ctx.to_device(g), ctx.to_device(m)
        ncc, mom = ctx.ncc(F, G, M, n)
        ctx.ncc_noalloc(F, G, M, n, out)
        got = out.download(10, np.float64)
        print(f"n={n}: asc_ncc {ncc!r}, noalloc {got[9]!r}")
        assert got[:9].tobytes() == mom.tobytes()
        assert got[9].tobytes() == np.float64(ncc).tobytes()
        # the moments are the data's (not a stale call's)
        assert got[3] == float((m > 0.5).sum())
        for a in (F, G, M):
            a.free()


def test_noalloc_ncc_refuses_oversize(asc, ctx):
    out = ctx.alloc(10 * 8)
    L = asc.lib()
    rc = L.asc_ncc_noalloc(ctx.handle, out.ptr, out.ptr, out.ptr, 9_000_001, out.ptr)
    assert rc == -5  # ASC_ERR_TOO_LARGE
    assert b"ASC_NCC_NOALLOC_MAX_N" in L.asc_last_error()
