"""Host-side check of the ray march's element addressing for every volume
layout (autoscoper_amd/csrc/device.hpp cell_offset() and brick_tables(),
kernels.hip pack kernels, runtime.hip layout_shape() and brick_view()),
emulated in integer arithmetic on small volumes of odd and even sizes: for
every cell the march can sample (voxel cell indices -1 .. n-1 per axis) the
byte offset is inside the array, the 24-bit multiplicands are in range, and
the element found there holds that cell's eight corner texels (zero outside
the volume). For each of BRICK's three copies (bricks across rays along x,
y or z) the pack places every covered cell at a distinct element and fills
the array exactly. The GPU tests then check that all layouts give
bit-identical DRRs.
"""
import itertools

import numpy as np
import pytest


def brick(W, H, D, axis, tex):
    """runtime.hip brick_view: the BRICK copy for rays along `axis` (bricks
    1 x 2 x 4 cells across the ray, the ray axis fastest) -> (E, blog, bin,
    bstr, nbytes)."""
    E = 8 * tex
    n = (W, H, D)
    u, w = (axis + 1) % 3, (axis + 2) % 3
    blog, bin_, bstr = [0] * 3, [0] * 3, [0] * 3
    blog[axis], blog[u], blog[w] = 0, 1, 2
    bin_[axis], bin_[u], bin_[w] = 0, E, 2 * E
    nb = [((n[a] - 1) >> blog[a]) + 2 for a in range(3)]
    bstr[axis] = 8 * E
    bstr[u] = nb[axis] * bstr[axis]
    bstr[w] = nb[u] * bstr[u]
    return E, blog, bin_, bstr, nb[w] * bstr[w], nb


def term(c, l, bi, bs):
    """One axis's offset term of a BRICK cell (device.hpp brick_tables)."""
    return bi * (c & ((1 << l) - 1)) + bs * ((c >> l) + 1)


def shape(W, H, D, layout, tex):
    if layout.startswith("brick"):
        E, blog, bin_, bstr, nbytes, nb = brick(W, H, D, int(layout[5]), tex)
        return E, 0, 0, nbytes
    per = {"linear": 1, "quad": 4, "octet": 8}[layout]
    esz = per * tex
    pW, pS = W + 2, (W + 2) * (H + 2)
    return esz, pW, pS, pS * (D + 2) * esz


def voxel(vol, x, y, z):
    D, H, W = vol.shape
    if 0 <= x < W and 0 <= y < H and 0 <= z < D:
        return int(vol[z, y, x])
    return 0


def packed_array(vol, layout, tex):
    """Element contents as the pack kernels write them: a dict offset -> corners."""
    D, H, W = vol.shape
    esz, pW, pS, nbytes = shape(W, H, D, layout, tex)
    mem = {}
    if layout.startswith("brick"):
        # kernels.hip pack_brick_kernel: every cell the bricks cover, placed
        # by its three terms; no two cells share an element
        E, blog, bin_, bstr, nbytes_b, nb = brick(W, H, D, int(layout[5]), tex)
        ranges = [range(-(1 << blog[a]), (nb[a] - 1) << blog[a]) for a in range(3)]
        for ck, cj, ci in itertools.product(ranges[2], ranges[1], ranges[0]):
            off = sum(term(c, blog[a], bin_[a], bstr[a]) for a, c in enumerate((ci, cj, ck)))
            assert off % E == 0 and 0 <= off <= nbytes_b - E and off not in mem
            mem[off] = tuple(voxel(vol, ci + (q & 1), cj + ((q >> 1) & 1), ck + (q >> 2)) for q in range(8))
        assert len(mem) * E == nbytes_b
    else:
        per = esz // tex
        for k, j, i in itertools.product(range(D + 2), range(H + 2), range(W + 2)):
            e = (k * (H + 2) + j) * (W + 2) + i
            mem[e * esz] = tuple(voxel(vol, i - 1 + (q & 1), j - 1 + ((q >> 1) & 1), k - 1 + (q >> 2))
                                 for q in range(per))
    return mem, esz, pW, pS, nbytes


def mad_i24(a, b, c):
    assert -(1 << 23) <= a < (1 << 23) and -(1 << 23) <= b < (1 << 23)
    return a * b + c


def offset(i, j, k, layout, esz, pW, pS, dims=None, tex=None):
    sy, sz = pW * esz, pS * esz
    if layout.startswith("brick"):
        # device.hpp cell_offset: the sum of the three LDS table entries
        W, H, D = dims
        assert W + H + D + 3 <= 4096  # kBrickTab
        E, blog, bin_, bstr, nbytes_b, nb = brick(W, H, D, int(layout[5]), tex)
        tab = [term(c, blog[0], bin_[0], bstr[0]) for c in range(-1, W)]
        tab += [term(c, blog[1], bin_[1], bstr[1]) for c in range(-1, H)]
        tab += [term(c, blog[2], bin_[2], bstr[2]) for c in range(-1, D)]
        return tab[i + 1] + tab[W + 2 + j] + tab[W + H + 3 + k]
    c0 = sy + sz + esz
    return mad_i24(k, sz, mad_i24(j, sy, i * esz + c0))


@pytest.mark.parametrize("layout", ["linear", "quad", "octet", "brick0", "brick1", "brick2"])
@pytest.mark.parametrize("W,H,D", [(1, 1, 1), (2, 3, 4), (5, 4, 3), (6, 7, 2)])
@pytest.mark.parametrize("tex", [2, 4])
def test_every_sampled_cell_is_addressed_in_bounds(layout, W, H, D, tex):
    rng = np.random.default_rng(W * 100 + H * 10 + D)
    vol = rng.integers(1, 1000, (D, H, W))
    mem, esz, pW, pS, nbytes = packed_array(vol, layout, tex)
    for k, j, i in itertools.product(range(-1, D), range(-1, H), range(-1, W)):
        off = offset(i, j, k, layout, esz, pW, pS, (W, H, D), tex)
        assert 0 <= off and off + esz <= nbytes, (i, j, k, off, nbytes)
        corners = mem[off]
        if layout == "linear":
            # LINEAR gathers x-pairs at +0, +sy, +sz, +sy+sz: check the first texel
            assert corners[0] == voxel(vol, i, j, k)
            assert off + 2 * esz + (pW + pS) * esz <= nbytes  # the x-pair at +sy+sz
        elif layout == "quad":
            assert corners == tuple(voxel(vol, i + (q & 1), j + (q >> 1), k) for q in range(4))
            assert off + pS * esz + esz <= nbytes  # the k+1 block
        else:
            assert corners == tuple(voxel(vol, i + (q & 1), j + ((q >> 1) & 1), k + (q >> 2)) for q in range(8))
