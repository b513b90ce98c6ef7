"""Host-side check of the ray march's element addressing for every volume
layout (autoscoper_amd/csrc/device.hpp trilinear(), kernels.hip pack
kernels, runtime.hip layout_shape()), emulated in integer arithmetic on
small volumes of odd and even sizes: for every cell the march can sample
(voxel cell indices -1 .. n-1 per axis) the byte offset is inside the
array, the 24-bit multiplicands are in range, and the element found there
holds that cell's eight corner texels (zero outside the volume). The GPU
tests then check that all layouts give bit-identical DRRs.
"""
import itertools

import numpy as np
import pytest


def shape(W, H, D, layout, tex):
    if layout == "brick":
        BX, BY, BZ = (W + 3) // 2, (H + 3) // 2, (D + 3) // 2
        esz = 8 * tex
        pW, pS = 8 * BX, 8 * BX * BY
        return esz, pW, pS, pS * BZ * esz
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
    if layout == "brick":
        BX, BY, BZ = (W + 3) // 2, (H + 3) // 2, (D + 3) // 2
        for bz, by, bx, c in itertools.product(range(BZ), range(BY), range(BX), range(8)):
            # cell c = (j&1) + 2(k&1) + 4(i&1) of the brick: x is the high bit
            ci, cj, ck = 2 * bx + (c >> 2) - 2, 2 * by + (c & 1) - 2, 2 * bz + ((c >> 1) & 1) - 2
            e = ((bz * BY + by) * BX + bx) * 8 + c
            mem[e * esz] = tuple(voxel(vol, ci + (q & 1), cj + ((q >> 1) & 1), ck + (q >> 2)) for q in range(8))
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


def offset(i, j, k, layout, esz, pW, pS):
    sy, sz = pW * esz, pS * esz
    if layout == "brick":
        c0 = sy + sz + 8 * esz
        E = esz
        base = i * 4 * E + j * E + k * 2 * E + c0
        return mad_i24(j >> 1, sy - 2 * E, mad_i24(k >> 1, sz - 4 * E, base))
    c0 = sy + sz + esz
    return mad_i24(k, sz, mad_i24(j, sy, i * esz + c0))


@pytest.mark.parametrize("layout", ["linear", "quad", "octet", "brick"])
@pytest.mark.parametrize("W,H,D", [(1, 1, 1), (2, 3, 4), (5, 4, 3), (6, 7, 2)])
@pytest.mark.parametrize("tex", [2, 4])
def test_every_sampled_cell_is_addressed_in_bounds(layout, W, H, D, tex):
    rng = np.random.default_rng(W * 100 + H * 10 + D)
    vol = rng.integers(1, 1000, (D, H, W))
    mem, esz, pW, pS, nbytes = packed_array(vol, layout, tex)
    for k, j, i in itertools.product(range(-1, D), range(-1, H), range(-1, W)):
        off = offset(i, j, k, layout, esz, pW, pS)
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
