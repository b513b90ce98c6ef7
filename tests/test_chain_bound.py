"""The filtered-DRR error bound used by the GPU parity tests holds: inputs
moved anywhere within the per-pixel DRR bar (random and worst-case signs)
move the oracle's filtered output by no more than the bound, for each
filter kind the DRR chains use. CPU only."""
import numpy as np
import pytest
from scipy import ndimage

from autoscoper_amd import filters as F
from chain_bound import chain_bound, drr_bar


@pytest.fixture(scope="module")
def drr():
    rng = np.random.default_rng(31)
    a = ndimage.gaussian_filter(rng.random((40, 56)), 2.0)
    a = (a - a.min()) / (a.max() - a.min()) * 0.9
    a[:6, :9] = 0  # a silhouette: exact zeros
    return a.astype(np.float32)


CHAINS = {
    "sobel": [F.sobel(1.7, 0.1)],
    "sobel-blend": [F.sobel(2.0, 0.7)],
    "gauss-sharpen": [F.gaussian(1.0), F.sobel(1.0, 0.5, enabled=False), F.sharpen(1.0, 2.0, 0.0)],
    "sharpen-thr": [F.sharpen(2.0, 0.5, 0.05)],
    "none": [],
}


@pytest.mark.parametrize("name", list(CHAINS))
def test_bound_holds_under_perturbation(orc, drr, name):
    recs = F.to_records(CHAINS[name])
    base = orc.filter_chain(recs, drr).astype(np.float64)
    bar = drr_bar(drr)
    bound = chain_bound(recs, drr, bar)
    rng = np.random.default_rng(7)
    worst = 0.0
    for k in range(40):
        if k < 20:
            d = rng.uniform(-1, 1, drr.shape) * bar
        else:  # all at the bar, random signs
            d = np.sign(rng.uniform(-1, 1, drr.shape)) * bar
        # 0.999: room for the float rounding of the perturbed input itself
        moved = orc.filter_chain(recs, (drr + 0.999 * d).astype(np.float32)).astype(np.float64)
        err = np.abs(moved - base)
        assert (err <= bound).all(), name
        worst = max(worst, float((err / np.maximum(bound, 1e-30)).max()))
    # the bound is not vacuous: it is reached to within a modest factor
    assert worst > 0.02 or name == "none", worst
    # (the threshold Sharpen's jump c*|f - blur| ~ c*thr dominates its bound)
    assert bound.max() < 0.1
