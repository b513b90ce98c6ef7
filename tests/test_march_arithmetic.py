"""Host-side check of the ray march's u16 trilinear arithmetic and texel
encodings (autoscoper_amd/csrc/device.hpp trilinear(), kernels.hip
store_element()), emulated with exact rational arithmetic and one rounding
to float32 per FMA, as the device's fmaf rounds:

- the three encodings of a 2 x 2 (x, y) block (plain, 15-bit and 14-bit
  differences) decode to the same integers x0, dx, dy, dxy;
- the slice form fma(b, fma(a, dxy, dy), fma(a, dx, x0)) and the z-lerp on
  two slices give a value >= 0 that is 0 exactly when every texel with a
  positive weight is 0 (the zero pattern of the specification's weighted
  sum, OpenCL 1.2 section 8.2, which decides the DRR's mask), and that lies
  within a few float32 ulps of the cell's largest texel of the exact
  weighted sum. (Relative to the value itself the error can be large where
  every large texel has a weight within an ulp of 0, as it is for nested
  lerps generally; a DRR pixel sums many samples.)
"""
from fractions import Fraction

import numpy as np
import pytest

F32 = np.float32


def round_f32(q: Fraction) -> float:
    """q rounded to the nearest float32, ties to even (q within range)."""
    x = F32(float(q))  # at most one float32 step from the correctly rounded value
    best = None
    for c in (np.nextafter(x, F32(-np.inf)), x, np.nextafter(x, F32(np.inf))):
        d = abs(Fraction(float(c)) - q)
        if best is None or d < best[0] or (d == best[0] and (int(c.view(np.uint32)) & 1) == 0):
            best = (d, c)
    return float(best[1])


def fma(a, b, c) -> float:
    return round_f32(Fraction(a) * Fraction(b) + Fraction(c))


def sub(a, b) -> float:
    return round_f32(Fraction(a) - Fraction(b))


def encode(block, enc):
    """kernels.hip store_element for one 2 x 2 block {t00, t10, t01, t11}."""
    t00, t10, t01, t11 = block
    if enc == 0:
        return [t00, t10, t01, t11]
    if enc == 1:
        return [t00, t10 - t00, t01 - t00, t11 - t01]
    return [t00, t10 - t00, t01 - t00, (t11 - t01) - (t10 - t00)]


def decode(e, enc):
    """device.hpp: the integers (x0, dx, dy, dxy) of one slice, as the march
    forms them from each encoding (16-bit halves sign-extended)."""
    s16 = [int(np.int16(np.uint16(v & 0xFFFF))) for v in e]
    if enc == 0:
        u = [v & 0xFFFF for v in e]
        return u[0], u[1] - u[0], u[2] - u[0], (u[3] - u[2]) - (u[1] - u[0])
    if enc == 1:
        return e[0] & 0xFFFF, s16[1], s16[2], s16[3] - s16[1]
    return e[0] & 0xFFFF, s16[1], s16[2], s16[3]


@pytest.mark.parametrize("enc,vmax", [(0, 65535), (1, 32767), (2, 16383)])
def test_encodings_decode_to_the_same_integers(enc, vmax):
    rng = np.random.default_rng(5 + enc)
    for _ in range(4000):
        block = [int(v) for v in rng.integers(0, vmax + 1, 4)]
        if rng.random() < 0.3:
            block[int(rng.integers(0, 4))] = vmax
        stored = [v & 0xFFFF for v in encode(block, enc)]  # 16-bit storage
        t00, t10, t01, t11 = block
        assert decode(stored, enc) == (t00, t10 - t00, t01 - t00, (t11 - t01) - (t10 - t00))


def trilinear_u16(t, a, b, c):
    """device.hpp trilinear() for u16 volumes; t = 8 corner texels
    {t000, t100, t010, t110, t001, t101, t011, t111}."""
    v = []
    for s in (0, 4):
        x0, dx, dy, dxy = t[s], t[s + 1] - t[s], t[s + 2] - t[s], (t[s + 3] - t[s + 2]) - (t[s + 1] - t[s])
        v.append(fma(b, fma(a, dxy, dy), fma(a, dx, x0)))
    return fma(c, sub(v[1], v[0]), v[0])


def weights(a, b, c):
    A, B, C = Fraction(a), Fraction(b), Fraction(c)
    return [(1 - A) * (1 - B) * (1 - C), A * (1 - B) * (1 - C), (1 - A) * B * (1 - C), A * B * (1 - C),
            (1 - A) * (1 - B) * C, A * (1 - B) * C, (1 - A) * B * C, A * B * C]


def test_slice_form_sign_zero_pattern_and_error():
    rng = np.random.default_rng(17)
    special = [0.0, float(np.nextafter(F32(1), F32(0))), float(F32(2.0 ** -24)), 0.5]
    worst = 0.0
    for k in range(6000):
        t = [int(v) for v in rng.integers(0, 16384, 8)]
        # silhouette-like cells: most corners zero
        for i in range(8):
            if rng.random() < 0.6:
                t[i] = 0
        w3 = [float(F32(rng.random())) if rng.random() < 0.8 else special[int(rng.integers(0, 4))] for _ in range(3)]
        a, b, c = w3
        got = trilinear_u16(t, a, b, c)
        wts = weights(a, b, c)
        exact = sum(wi * ti for wi, ti in zip(wts, t))
        assert got >= 0.0
        assert (got == 0.0) == (exact == 0), (t, w3, got)
        if exact:
            err = abs(Fraction(got) - exact) / max(t)
            worst = max(worst, float(err))
    # a handful of roundings of values up to the cell's largest texel
    assert worst <= 8 * 2.0 ** -24, worst
