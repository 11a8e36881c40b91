"""CPU emulation of the step and imaging kernels' index arithmetic.

Mirrors tile_range / window_start of fwi2d_amd/csrc/kernels.hip and checks, for many
grid shapes and every tile, that each stencil read of the kernels (SBP interior and
closure rows, first-derivative products, SAT rows, extraction footprints) lands inside
the LDS arrays and on a cell of the grid that the tile loaded -- so the kernels never
read outside their shared-memory windows nor use a zero-filled halo cell as data.
Also checks that the tiles' write ranges partition the grid.
"""
import numpy as np
import pytest

TZ, H_STEP, H_IMG = 64, 4, 2


def tile_range(k, ntile, T, n):
    if ntile == 1:
        return 0, n
    if k == ntile - 1:
        return n - T, n
    s = k * T
    return s, min(s + T, n - T)


def window_start(s, H):
    return 0 if (H <= s < H + 2) else s - H


def d1_offsets(g, n):
    if 4 <= g < n - 4:
        return [-2, -1, 1, 2]
    if g < 4:
        return [j - g for j in range(6)]
    r = n - 1 - g
    return [r - j for j in range(6)]


def d2_offsets(g, n):
    if 6 <= g < n - 6:
        return [-2, -1, 0, 1, 2]
    if g < 6:
        return [j - g for j in range(8)]
    r = n - 1 - g
    return [r - j for j in range(8)]


def needed_products(s, e, n):
    """Range of first-derivative products a tile writing [s, e) uses (mirrors the kernel)."""
    lo, hi = max(0, s - 2), min(n, e + 2)
    if s < 4:
        lo, hi = 0, max(hi, 6)
    if e > n - 4:
        lo, hi = min(lo, n - 6), n
    return lo, hi


def tiles(n, T):
    nt = (n + T - 1) // T
    return [tile_range(k, nt, T, n) for k in range(nt)]


def tile_range_quads(k, ntile, T, n):
    """z tiling in 4-cell quads: tiles start on quads; the last one is anchored to the
    last quad of the padded extent ceil(n/4)*4 and writes up to n."""
    nq, tq = (n + 3) // 4, T // 4
    if ntile == 1:
        return 0, n
    if k == ntile - 1:
        return 4 * (nq - tq), n
    s = k * T
    return s, min(s + T, 4 * (nq - tq))


def tiles_quads(n, T):
    nq, tq = (n + 3) // 4, T // 4
    nt = (nq + tq - 1) // tq
    return [tile_range_quads(k, nt, T, n) for k in range(nt)]


def check_axis(n, T, quads=False):
    """Per-axis check (the 2-D stencils are separable, so axes can be checked apart)."""
    W = T + 2 * H_STEP
    I = T + 8 if quads else T + 4
    cover = np.zeros(n, int)
    for s, e in (tiles_quads(n, T) if quads else tiles(n, T)):
        if quads:
            assert s % 4 == 0
        assert 0 <= s < e <= n and e - s <= T
        cover[s:e] += 1
        ws = window_start(s, H_STEP)
        if quads:
            assert ws % 4 == 0
        loaded = set(p for p in range(ws, ws + W) if 0 <= p < n)
        # first-derivative products: computed on the needed range [plo, phi), stored from plo
        plo, phi = needed_products(s, e, n)
        inter = set(range(plo, phi))
        assert phi - plo <= I, (n, T, s, plo, phi)
        porg = ws if quads else plo  # quad kernels store products on the state window
        assert inter <= loaded  # cross-axis reads of a product use its own row/column
        for p in inter:
            assert 0 <= p - porg < I
            for o in d1_offsets(p, n):
                assert 0 <= p + o - ws < W and (p + o) in loaded, (n, T, s, p, o)
        for g in range(s, e):
            for o in d2_offsets(g, n):
                assert 0 <= g + o - ws < W and (g + o) in loaded, (n, T, s, g, o)
            for o in d1_offsets(g, n):
                # D1 of the products along this axis and of the state (tangential SAT rows)
                assert 0 <= g + o - porg < I and (g + o) in inter, (n, T, s, g, o)
                assert (g + o) in loaded
            # SAT normal rows (4 one-sided points, 6 closure points) at the edges
            if g == 0:
                assert all(k in loaded for k in range(6))
            if g == n - 1:
                assert all(n - 1 - k in loaded for k in range(6))
            # extraction footprint of a pivot owned by this tile
            if 1 <= g <= n - 2:
                assert all((g + o) in loaded and 0 <= g + o - ws < W for o in (-1, 0, 1))
        # imaging kernel (halo 2)
        wsi = window_start(s, H_IMG)
        Wi = T + 2 * H_IMG
        loaded_i = set(p for p in range(wsi, wsi + Wi) if 0 <= p < n)
        for g in range(s, e):
            for o in d1_offsets(g, n):
                assert 0 <= g + o - wsi < Wi and (g + o) in loaded_i, ("image", n, T, s, g, o)
    assert np.all(cover == 1), (n, T, cover)


@pytest.mark.parametrize("T", [8, 16, 32, 64])
def test_tiles_partition_and_windows(T):
    for n in list(range(12, 3 * T + 20)) + [201, 300, 500, 601, 1000, 1700, 3000]:
        check_axis(n, T)


def test_quad_tiles_along_z():
    for n in list(range(12, 300)) + [501, 601, 1000, 3001]:
        check_axis(n, 64, quads=True)


def xcd_remap(bid, nb):
    q, r = nb >> 3, nb & 7
    xcd, local = bid & 7, bid >> 3
    return xcd * (q + 1) + local if xcd < r else r * (q + 1) + (xcd - r) * q + local


def test_xcd_remap_is_a_permutation():
    for nb in list(range(1, 300)) + [4096, 27001]:
        m = [xcd_remap(b, nb) for b in range(nb)]
        assert sorted(m) == list(range(nb))
