"""How far a filtered DRR may move when its input DRR moves within the
pixel bar.

The DRR bar is per pixel: |gpu - oracle| <= 1e-4 * |oracle| (plus, with a
cutoff, the contribution of samples near the cutoff, which a sampler
rounding differently may flip; see oracle.drr_cut_margin). A filter turns
an input error into an output error; these are the bounds, stage by stage,
in exact arithmetic, plus a rounding allowance of the stage's own float
evaluation (a few ulps of the magnitudes it adds):

  Sobel      out = clamp(c1*|(H, V)| + c2*p11), H and V sums of 8 unit
             weights over the clamped 3x3 window, so
             |d out| <= c1*sqrt(EH^2 + EV^2) + c2*e11, with EH, EV the
             weighted sums of the input bounds (clamp is 1-Lipschitz);
             c1 = scale, c2 = 2*blend when blend < 1/2, else
             c1 = 2*(1-blend)*scale, c2 = 1 (SobelFilter.cl:36-44).
  Gaussian   a zero-padded convolution with non-negative weights, then a
             clamp: |d out| <= sum w*e.
  Sharpen    blur as Gaussian without the clamp; out = clamp(f + c*(f - blur))
             where |f - blur| > thr, else f. Each branch is Lipschitz
             ((1+c)*e + c*e_blur); where |f - blur| lies within e + e_blur
             of thr the branch may differ and the jump c*|f - blur| is added.
  disabled   identity.
  Contrast   not bounded here (pow near zero); not used on DRR chains.

Sizes are small (test images), so plain numpy is enough.
"""
import numpy as np

import oracle as O

F_CONTRAST, F_SOBEL, F_MEDIAN, F_GAUSSIAN, F_SHARPEN = 0, 1, 2, 3, 4
ULP = 2.0 ** -24


def _clamped_window(a):
    """The 3x3 neighbourhood of every pixel, edge-clamped: (9, h, w) in
    row-major window order p00 .. p22."""
    p = np.pad(a, 1, mode="edge")
    h, w = a.shape
    return np.stack([p[r:r + h, c:c + w] for r in range(3) for c in range(3)])


def _conv_zero(a, wts):
    """sum over in-image taps of w[i][j] * a[y - R + j][x - R + i] (the
    reference's weight indexing), zero outside."""
    n = wts.shape[0]
    R = (n - 1) // 2
    h, w = a.shape
    p = np.pad(a, R)
    out = np.zeros_like(a, dtype=np.float64)
    for i in range(n):
        for j in range(n):
            out += wts[i, j] * p[j:j + h, i:i + w]
    return out


def stage_bound(rec, f, e):
    """Bound on |filter(f + d) - filter(f)| for |d| <= e, one record."""
    kind, enabled = int(rec[0]), rec[1] != 0
    f = f.astype(np.float64)
    if not enabled:
        return e.copy()
    if kind == F_SOBEL:
        scale, blend = float(rec[2]), float(rec[3])
        c1, c2 = (scale, 2 * blend) if blend < 0.5 else (2 * (1 - blend) * scale, 1.0)
        E = _clamped_window(e)
        P = np.abs(_clamped_window(f))
        hw = np.array([1, 0, 1, 2, 0, 2, 1, 0, 1], float)[:, None, None]  # |coefficients| of H
        vw = np.array([1, 2, 1, 0, 0, 0, 1, 2, 1], float)[:, None, None]
        EH, EV = (hw * E).sum(0), (vw * E).sum(0)
        rnd = 16 * ULP * (c1 * 8 * P.max(0) + c2 * np.abs(f))
        return c1 * np.sqrt(EH ** 2 + EV ** 2) + c2 * e + rnd
    if kind in (F_GAUSSIAN, F_SHARPEN):
        wts = O.gaussian_weights(float(rec[2])).astype(np.float64)
        if wts.shape[0] == 1:
            return e.copy() if kind == F_GAUSSIAN else e.copy()
        eb = _conv_zero(e, wts) + 16 * ULP * _conv_zero(np.abs(f), wts)
        if kind == F_GAUSSIAN:
            return eb
        c, thr = max(float(rec[3]), 1.0), float(rec[4])
        blur = _conv_zero(f, wts)
        d = np.abs(f - blur)
        out = (1 + c) * e + c * eb + 16 * ULP * (1 + c) * (np.abs(f) + np.abs(blur))
        near = np.abs(d - thr) <= e + eb + 16 * ULP * (np.abs(f) + np.abs(blur))
        return np.where(near, out + c * (d + e + eb), out)
    raise NotImplementedError(f"no error bound for filter type {kind}")


def chain_bound(records, drr, e_in):
    """Per-pixel bound on the filtered DRR's error, given the input DRR (the
    oracle's) and the bound on its error."""
    recs = np.asarray(records, np.float32).reshape(-1, 6)
    f = drr.astype(np.float32)
    e = np.asarray(e_in, np.float64)
    for rec in recs:
        e = stage_bound(rec, f, e)
        f = O.filter_chain(rec[None], f)
    return e


def drr_bar(drr, margin=None):
    """The DRR pixel bar: 1e-4 relative, plus the cutoff margin if any."""
    e = 1e-4 * np.abs(drr.astype(np.float64))
    return e if margin is None else e + margin


def assert_within(got, want, bound, what=""):
    err = np.abs(got.astype(np.float64) - want.astype(np.float64))
    bad = err > bound
    assert not bad.any(), (f"{what}: {int(bad.sum())} pixels beyond the propagated bound; worst excess "
                           f"{float((err - bound).max()):.3e}")
