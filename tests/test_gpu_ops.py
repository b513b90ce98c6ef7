"""Per-operator HIP kernels through the C ABI against the float-faithful
oracle on the same inputs.

Tolerances (stated per the parity bar):
  DRR pixels            |gpu - oracle| <= 1e-4 * |oracle|, zero pattern exact;
                        with a cutoff, plus the oracle's cutoff margin
  radiograph/background bit-exact (same float operation order)
  Sobel/Gaussian/Sharpen bit-exact; Contrast <= 4 ulp-scale (powf)
  NCC                   |gpu - oracle FP64| <= 1e-5, mask count exact
"""
import numpy as np
import pytest

from autoscoper_amd import synth

pytestmark = pytest.mark.gpu


def assert_drr_close(g, o):
    assert g.shape == o.shape
    assert ((g == 0) == (o == 0)).all(), "DRR zero pattern differs"
    nz = o != 0
    rel = np.abs(g[nz].astype(np.float64) - o[nz]) / np.abs(o[nz])
    assert rel.max(initial=0) <= 1e-4, rel.max()


@pytest.fixture(scope="module")
def vol_pair(asc, ctx, tiny):
    v = asc.Volume(ctx, tiny["wl"].volume, scale=(tiny["wl"].voxel_mm,) * 3)
    yield v, tiny["prep"]
    v.close()


def test_volume_description_matches_oracle(asc, ctx, orc):
    rng = np.random.default_rng(3)
    data = np.zeros((9, 11, 13), np.uint16)
    data[2:7, 3:9, 1:12] = rng.integers(1, 60000, (5, 6, 11))
    data[4, 5, 6] = 0
    for flip in [(0, 0, 0), (1, 0, 1), (0, 1, 0)]:
        v = asc.Volume(ctx, data, scale=(0.3, 0.4, 0.5), flip=flip)
        ref = orc.volume_prep(data, (0.3, 0.4, 0.5), flip)
        d = v.desc
        assert list(d.dims) == list(ref["dims"])
        assert list(d.crop_min) == list(ref["crop_min"]) and list(d.crop_max) == list(ref["crop_max"])
        assert np.array_equal(np.array(d.inv_scale, np.float32), ref["inv_scale"])
        assert np.array_equal(np.array(d.inv_trans, np.float32), ref["inv_trans"])
        assert np.array_equal(np.array(d.trans_center), ref["trans_center"])
        assert np.float32(d.min_value) == np.float32(ref["min_value"])
        assert np.float32(d.max_value) == np.float32(ref["max_value"])
        v.close()
    with pytest.raises(asc.AscError):
        asc.Volume(ctx, np.zeros((4, 4, 4), np.uint16))


@pytest.mark.parametrize("storage", [0, 3])
def test_drr_matches_oracle(asc, ctx, orc, tiny, storage):
    wl, prep = tiny["wl"], tiny["prep"]
    v = asc.Volume(ctx, wl.volume, scale=(wl.voxel_mm,) * 3, storage=storage)
    RW, RH = tiny["render"]
    rng = np.random.default_rng(11)
    poses = wl.gt_poses[0] + np.concatenate([rng.uniform(-4, 4, (6, 3)), rng.uniform(-15, 15, (6, 3))], 1)
    total = 0
    for cam in tiny["cams"]:
        for p in poses:
            inview, imv, vp, (w, h) = orc.view_geometry(cam, p, prep["inv_scale"], prep["inv_trans"], RW, RH)
            if not inview:
                continue
            ref, n = orc.drr(prep["volume"], imv[:12], vp, synth.DRR_STEP, synth.DRR_INTENSITY, 0.0, w, h)
            out = ctx.alloc(4 * w * h)
            ctx.drr_render(v, imv[:12], vp, synth.DRR_STEP, synth.DRR_INTENSITY, 0.0, out, w, h)
            assert_drr_close(out.download((h, w)), ref)
            total += n
    assert total > 0
    v.close()


@pytest.mark.parametrize("storage,gain", [(0, 1), (0, 8), (0, 16), (1, 1), (2, 1), (3, 1)])
def test_layouts_are_bit_identical(asc, ctx, orc, tiny, storage, gain):
    """QUAD/OCTET/BRICK corner-replicated layouts change only the access
    pattern. The paired layouts of a u16 volume store differences of its
    texels: the row differences too when every value is <= 16383 (gain 1),
    the x-differences alone up to 32767 (gain 8), and the texels as they are
    above that (gain 16). Every encoding gives the LINEAR layout's bits."""
    wl, prep = tiny["wl"], tiny["prep"]
    data = np.minimum(wl.volume.astype(np.uint32) * gain, 65535).astype(np.uint16)
    assert (data.max() <= 16383) == (gain == 1) and (data.max() > 32767) == (gain == 16)
    v = asc.Volume(ctx, data, scale=(wl.voxel_mm,) * 3, storage=storage)
    RW, RH = tiny["render"]
    out = ctx.alloc(4 * RW * RH)
    # camera 0 at the hidden pose looks along x; pitched 90 degrees, along z;
    # camera 1 looks along y: each of BRICK's three copies is read
    W, H, D = prep["dims"]
    axes = set()
    for cam, turn in ((tiny["cams"][0], (0, 0, 0)), (tiny["cams"][0], (0, 90, 0)), (tiny["cams"][1], (0, 0, 0))):
        pose = wl.gt_poses[0] + np.array([0, 0, 0, *turn])
        _, imv, _, _ = orc.view_geometry(cam, pose, prep["inv_scale"], prep["inv_trans"], RW, RH)
        axes.add(int(np.argmax([abs(imv[2]) * W, abs(imv[6]) * H, abs(imv[10]) * D])))
        fvp = synth.full_frame_viewport(cam[24:28], RW, RH)
        imgs = []
        for layout in (v.LINEAR, v.QUAD, v.OCTET, v.BRICK, v.LINEAR):
            v.set_layout(layout)
            ctx.drr_render(v, imv[:12], fvp, synth.DRR_STEP, synth.DRR_INTENSITY, 0.0, out, RW, RH)
            imgs.append(out.download((RH, RW)))
        assert imgs[0].max() > 0
        for im in imgs[1:]:
            assert np.array_equal(im, imgs[0])
    assert axes == {0, 1, 2}
    v.close()


def test_drr_cutoff_and_full_frame(asc, ctx, orc, tiny, vol_pair):
    v, prep = vol_pair
    RW, RH = tiny["render"]
    cam = tiny["cams"][1]
    _, imv, _, _ = orc.view_geometry(cam, tiny["wl"].gt_poses[0], prep["inv_scale"], prep["inv_trans"], RW, RH)
    fvp = synth.full_frame_viewport(cam[24:28], RW, RH)
    cutoff = 0.5 * (prep["min_value"] + prep["max_value"])
    for c in (0.0, cutoff):
        ref, _ = orc.drr(prep["volume"], imv[:12], fvp, 0.37, 3.0, c, RW, RH)
        out = ctx.alloc(4 * RW * RH)
        ctx.drr_render(v, imv[:12], fvp, 0.37, 3.0, c, out, RW, RH)
        g = out.download((RH, RW))
        if c == 0.0:
            assert_drr_close(g, ref)
        else:
            # the sampler's rounding is implementation-defined, so a sample
            # within a few ulps of the cutoff may be kept by one evaluation
            # and dropped by another: beyond the 1e-4 bar a pixel may move by
            # at most the contribution of such samples (the oracle's margin)
            ref_m, margin = orc.drr_cut_margin(prep["volume"], imv[:12], fvp, 0.37, 3.0, c, RW, RH)
            assert np.array_equal(ref_m, ref)
            err = np.abs(g.astype(np.float64) - ref)
            assert (err <= 1e-4 * np.abs(ref) + margin).all(), float((err - 1e-4 * np.abs(ref) - margin).max())
            assert (margin > 0).mean() < 0.05 and (g > 0).mean() > 0.1


@pytest.mark.parametrize("kind", ["u8", "u16", "f32"])
def test_rad_and_background_bit_exact(asc, ctx, orc, kind):
    rng = np.random.default_rng(5)
    h, w = 77, 93
    base = rng.random((h, w)).astype(np.float32)
    img = {"u8": (base * 255).astype(np.uint8), "u16": (base * 65535).astype(np.uint16), "f32": base}[kind]
    im = asc.Image(ctx, img)
    plane = np.array([0.01, -0.02, 0.31, 0.27], np.float32)
    for vp in ([-0.1, -0.12, 0.2, 0.25], [-0.2, -0.2, 0.5, 0.45], [0.05, 0.0, 0.1, 0.1]):
        vp = np.array(vp, np.float32)
        ow, oh = 61, 53
        ref = orc.rad(img, plane, vp, ow, oh)
        out = ctx.alloc(4 * ow * oh)
        ctx.rad_render(im, plane, vp, out, ow, oh)
        assert np.array_equal(out.download((oh, ow)).view(np.uint32), ref.view(np.uint32))
    if kind == "f32":
        ref = orc.background(img, plane, vp, 0.4, ow, oh)
        ctx.background_render(im, plane, vp, 0.4, out, ow, oh)
        assert np.array_equal(out.download((oh, ow)), ref)
    im.close()


@pytest.fixture(scope="module")
def images():
    rng = np.random.default_rng(9)
    a = rng.random((45, 67)).astype(np.float32)
    a[:5, :7] = 0.0
    b = np.clip(0.5 + 0.2 * rng.standard_normal((31, 130)), 0, 1).astype(np.float32)
    return [a, b, np.full((4, 3), 0.25, np.float32), rng.random((1, 17)).astype(np.float32)]


def _run_filter(asc, ctx, f, img):
    h, w = img.shape
    src, dst = ctx.to_device(img), ctx.alloc(img.nbytes)
    ctx.filter_apply(f, src, dst, w, h)
    return dst.download((h, w))


def test_sobel_bit_exact(asc, ctx, orc, images):
    for img in images:
        for scale, blend in ((1.0, 0.5), (1.7, 0.1), (3.0, 0.43), (0.6, 0.9)):
            g = _run_filter(asc, ctx, asc.filters.sobel(scale, blend), img)
            assert np.array_equal(g.view(np.uint32), orc.sobel(img, scale, blend).view(np.uint32))


def test_contrast_close(asc, ctx, orc, images):
    for img in images:
        for alpha, beta in ((1.0, 1.0), (2.5, 0.9), (1.3, 2.0)):
            g = _run_filter(asc, ctx, asc.filters.contrast(alpha, beta, 3), img)
            o = orc.contrast(img, alpha, beta, 3)
            assert np.allclose(g, o, rtol=4e-7 * 8, atol=1e-30)
            assert ((g == 0) == (o == 0)).all()


def test_gaussian_sharpen_bit_exact(asc, ctx, orc, images):
    for img in images:
        for r in (0.0, 0.5, 1.0, 1.5, 3.0, 4.0):
            g = _run_filter(asc, ctx, asc.filters.gaussian(r), img)
            assert np.array_equal(g.view(np.uint32), orc.gaussian(img, r).view(np.uint32))
        for r, c, t in ((1.0, 1.0, 0.0), (1.0, 2.5, 0.05), (2.0, 0.5, 0.0)):
            g = _run_filter(asc, ctx, asc.filters.sharpen(r, c, t), img)
            assert np.array_equal(g.view(np.uint32), orc.sharpen(img, r, c, t).view(np.uint32))


def test_filter_chain_and_disabled(asc, ctx, orc, images):
    img = images[1]
    h, w = img.shape
    chains = [
        [],
        [asc.filters.sobel(3.0, 0.43)],
        [asc.filters.sobel(3.0, 0.43), asc.filters.contrast(2.5, 0.9)],
        [asc.filters.gaussian(1.0), asc.filters.sobel(1.0, 0.5, enabled=False), asc.filters.sharpen(1.0, 2.0, 0.0)],
    ]
    for ch in chains:
        src, dst, scr = ctx.to_device(img), ctx.alloc(img.nbytes), ctx.alloc(img.nbytes)
        ctx.filter_chain(ch, src, dst, scr, w, h)
        g = dst.download((h, w))
        o = orc.filter_chain(asc.filters.to_records(ch), img)
        assert np.allclose(g, o, rtol=1e-5, atol=1e-7)


def test_elementwise_ops(asc, ctx, orc):
    rng = np.random.default_rng(2)
    h, w = 19, 23
    a = rng.random((h, w)).astype(np.float32)
    a[rng.random((h, w)) < 0.3] = 0.0
    b = rng.random((h, w)).astype(np.float32)
    m = (rng.random((h, w)) < 0.5).astype(np.float32)
    da, db, dm, out = ctx.to_device(a), ctx.to_device(b), ctx.to_device(m), ctx.alloc(a.nbytes)
    ctx.multiply(da, db, out, w, h)
    assert np.array_equal(out.download((h, w)), a * b)
    ctx.merge(da, db, out, w, h)
    assert np.array_equal(out.download((h, w)), orc.merge(a, b))
    ctx.drr_background(da, out, w, h)
    assert np.array_equal(out.download((h, w)), (a != 0).astype(np.float32))
    comp = ctx.alloc(3 * a.nbytes)
    ctx.composite(da, db, dm, None, comp, w, h)
    assert np.array_equal(comp.download((h, w, 3)), orc.composite(a, b, m))
    # fused mask + multiplies (Tracker.cpp:747-752)
    mk = ctx.alloc(a.nbytes)
    ctx.mask_multiply(da, db, dm, mk, w, h)
    mm = (a != 0).astype(np.float32) * m
    assert np.array_equal(mk.download((h, w)), mm)
    assert np.array_equal(da.download((h, w)), a * mm)
    assert np.array_equal(db.download((h, w)), b * mm)


@pytest.mark.parametrize("n", [1, 100, 2047, 2048, 2049, 16384, 30001, 512 * 512 + 5])
def test_ncc_matches_fp64(asc, ctx, orc, n):
    rng = np.random.default_rng(n)
    f = rng.random(n).astype(np.float32)
    g = (0.3 * f + 0.7 * rng.random(n)).astype(np.float32)
    m = (rng.random(n) < 0.8).astype(np.float32)
    m[0] = 1.0
    f, g = f * m, g * m
    v, mom = ctx.ncc(ctx.to_device(f), ctx.to_device(g), ctx.to_device(m), n)
    ref, rmom = orc.ncc_exact(f.astype(np.float64), g.astype(np.float64), m.astype(np.float64))
    assert abs(v - ref) <= 1e-5 or (ref == 1e5 and v == 1e5)
    assert mom[3] == rmom[6]  # count(mask > 0.5), exact
    if n > 2:
        # independent check: numpy Pearson correlation over the mask
        sel = m > 0.5
        assert abs(v - np.corrcoef(f[sel], g[sel])[0, 1]) <= 1e-5


def test_ncc_degenerate(asc, ctx):
    n = 1000
    f = np.full(n, 0.5, np.float32)
    g = np.linspace(0, 1, n).astype(np.float32)
    m = np.ones(n, np.float32)
    v, _ = ctx.ncc(ctx.to_device(f), ctx.to_device(g), ctx.to_device(m), n)
    assert v == 1e5  # den < 1e-5 (Ncc.cpp:215-217)
    v, _ = ctx.ncc(ctx.to_device(g), ctx.to_device(2 * g + 0.1), ctx.to_device(m), n)
    assert abs(v - 1.0) <= 1e-6


@pytest.mark.parametrize("n", [1, 10, 100, 1000, 16384, 65536, 163841])
def test_ncc_reference_tree_forms_bit_exact(asc, ctx, orc, n):
    """asc_ncc in the reference's float form (six tree sums, each read back)
    equals the oracle's restatement of NccSum.cl / the CUDA sum kernel bit
    for bit, for both index models; sizes include non-power-of-two groups
    (n = 10, 100 drop elements in the LDS tree) and multi-pass sums."""
    rng = np.random.default_rng(n + 7)
    f = rng.random(n).astype(np.float32)
    g = (0.3 * f + 0.7 * rng.random(n)).astype(np.float32)
    m = (rng.random(n) < 0.8).astype(np.float32)
    m[0] = 1.0
    f, g = f * m, g * m
    df, dg, dm = ctx.to_device(f), ctx.to_device(g), ctx.to_device(m)
    try:
        for form, mode in ((asc.NCC_TREE_CUDA, 0), (asc.NCC_TREE_OPENCL, 1)):
            ctx.set_ncc_form(form)
            v, mom = ctx.ncc(df, dg, dm, n)
            ref = orc.ncc_tree(f, g, m, mode)
            assert v == ref or (np.isnan(v) and np.isnan(ref)), (form, v, ref)
            assert mom[0] == orc.tree_sum(m, mode) and mom[1] == orc.tree_sum(f, mode)
            assert np.isnan(mom[6:]).all()
        # a constant image under a full mask: whatever elements the tree
        # keeps, the mean is 0.5 exactly, den = 0 < 1e-5, 1e5, and the
        # products sum is skipped
        ctx.set_ncc_form(asc.NCC_TREE_OPENCL)
        c = ctx.to_device(np.full(n, 0.5, np.float32))
        ones = np.ones(n, np.float32)
        v, mom = ctx.ncc(c, dg, ctx.to_device(ones), n)
        assert v == 1e5 == orc.ncc_tree(np.full(n, 0.5, np.float32), g, ones, 1) and np.isnan(mom[3])
    finally:
        ctx.set_ncc_form(asc.NCC_MOMENTS)
    v, _ = ctx.ncc(df, dg, dm, n)
    ref, _ = orc.ncc_exact(f.astype(np.float64), g.astype(np.float64), m.astype(np.float64))
    assert abs(v - ref) <= 1e-5 or (ref == 1e5 and v == 1e5)  # the default form is back


def test_wide_slices_take_a_layout_that_fits(asc, ctx):
    """A 1100 x 1100 slice is too wide for OCTET/QUAD addressing (>= 8 MiB
    per slice step): asking for those is an error, not a wrong result.
    BRICK, addressed through its offset tables, takes it (creation picks
    it) and renders the LINEAR layout's bits."""
    vol = np.full((2, 1100, 1100), 3000, np.uint16)
    vol[:, 400:700, 300:350] = 1000
    v = asc.Volume(ctx, vol, scale=(0.5, 0.5, 0.5))
    for layout in (v.QUAD, v.OCTET):
        with pytest.raises(asc.AscError):
            v.set_layout(layout)
    # rays from above the unit box ([0, 1] x [0, 1] x [-1, 0]) down through
    # the middle of the slab
    imv = np.zeros(12, np.float32)
    imv[0], imv[5], imv[10] = 1.0, -1.0, 1.0
    imv[3], imv[7], imv[11] = 0.5, 0.5, 0.5
    vp = np.array([-0.3, -0.3, 0.6, 0.6])
    out = ctx.alloc(4 * 64 * 64)
    imgs = []
    for layout in (v.BRICK, v.LINEAR):
        v.set_layout(layout)
        ctx.drr_render(v, imv, vp, 0.01, 1.0, 0.0, out, 64, 64)
        imgs.append(out.download((64, 64)))
    assert imgs[0].max() > 0 and np.array_equal(imgs[0], imgs[1])
    v.close()

