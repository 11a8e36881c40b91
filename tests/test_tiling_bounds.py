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


def check_x_axis(n, T):
    """x axis of the step kernel (cells): window halo 4, products on [plo, phi) stored
    from plo in T + 4 columns; interior/closure choice per column."""
    W, I = T + 2 * H_STEP, T + 4
    cover = np.zeros(n, int)
    for s, e in tiles(n, T):
        assert 0 <= s < e <= n and e - s <= T
        cover[s:e] += 1
        ws = window_start(s, H_STEP)
        loaded = set(p for p in range(ws, ws + W) if 0 <= p < n)
        plo, phi = needed_products(s, e, n)
        need = set(range(plo, phi))
        assert phi - plo <= I, (n, T, s, plo, phi)
        for p in need:  # product columns: Dx of the state (fast: +-1, +-2; closure: 6 points)
            taps = [-2, -1, 1, 2] if 4 <= p < n - 4 else d1_offsets(p, n)
            for o in taps:
                assert 0 <= p + o - ws < W and (p + o) in loaded, (n, T, s, p, o)
        for g in range(s, e):
            d2 = [-2, -1, 0, 1, 2] if 6 <= g < n - 6 else d2_offsets(g, n)
            for o in d2:
                assert 0 <= g + o - ws < W and (g + o) in loaded, (n, T, s, g, o)
            for o in d1_offsets(g, n):  # Dx of the products, tangential SAT derivative
                assert 0 <= g + o - plo < I and (g + o) in need, (n, T, s, g, o)
                assert (g + o) in loaded
            if g == 0:
                assert all(k in loaded for k in range(6))
            if g == n - 1:
                assert all(n - 1 - k in loaded for k in range(6))
            if 1 <= g <= n - 2:  # extraction footprint of a pivot owned by this tile
                assert all((g + o) in loaded and 0 <= g + o - ws < W for o in (-1, 0, 1))
        check_image_axis(n, T, s, e)
    assert np.all(cover == 1), (n, T, cover)


HZ_STEP, PZ = 8, TZ + 8


def check_z_axis(n):
    """z axis of the step kernel (quads of 4 cells): window rows [s-8, s+72), product
    rows [s-4, s+68) in 18 quads; each quad takes the interior path only when all its
    cells are interior, else per-cell closure code."""
    ldz = (n + 3) // 4 * 4
    WZ = TZ + 2 * HZ_STEP
    cover = np.zeros(n, int)
    for s, e in tiles_quads(n, TZ):
        assert s % 4 == 0 and 0 <= s < e <= n and e - s <= TZ
        cover[s:e] += 1
        ws, pz0 = s - HZ_STEP, s - 4
        assert ws + WZ >= min(ldz, s + TZ + 8)
        loaded = set(p for p in range(ws, ws + WZ) if 0 <= p < n)  # padding rows are not data
        plo, phi = needed_products(s, e, n)
        need = set(range(plo, phi))
        assert plo >= pz0 and phi <= pz0 + PZ, (n, s, plo, phi)
        for lq in range(PZ // 4):
            z0 = pz0 + 4 * lq
            if not (z0 + 3 >= plo and z0 < phi):
                continue
            assert 0 <= z0 and z0 + 3 < ldz  # coefficient float4 read stays in the column
            if z0 >= 4 and z0 + 3 < n - 4:
                assert ws <= z0 - 4 and z0 + 7 < ws + WZ  # 12-row LDS read
                for k in range(4):
                    if z0 + k in need:
                        assert all(z0 + k + o in loaded for o in (-2, -1, 1, 2))
            else:
                for k in range(4):
                    g = z0 + k
                    if g < 0 or g >= n:
                        continue
                    for o in d1_offsets(g, n):
                        assert 0 <= g + o - ws < WZ and (g + o) in loaded, (n, s, g, o)
        for j in range(TZ // 4):
            z0 = s + 4 * j
            if z0 >= e:
                continue
            assert z0 + 3 < ldz
            if z0 >= 6 and z0 + 3 < n - 6:
                assert ws <= z0 - 4 and z0 + 7 < ws + WZ
                assert 0 <= z0 - 4 - pz0 and z0 + 7 - pz0 < PZ
                for k in range(4):
                    g = z0 + k
                    assert all(g + o in loaded for o in (-2, -1, 0, 1, 2))
                    assert all(g + o in need for o in (-2, -1, 1, 2))
            else:
                for k in range(4):
                    g = z0 + k
                    if g >= n:
                        continue
                    for o in d2_offsets(g, n):
                        assert 0 <= g + o - ws < WZ and (g + o) in loaded, (n, s, g, o)
                    for o in d1_offsets(g, n):
                        assert 0 <= g + o - pz0 < PZ and (g + o) in need, (n, s, g, o)
        for g in range(s, e):
            if g == 0:
                assert all(k in loaded for k in range(6))
            if g == n - 1:
                assert all(n - 1 - k in loaded for k in range(6))
            if 1 <= g <= n - 2:
                assert all((g + o) in loaded and 0 <= g + o - ws < WZ for o in (-1, 0, 1))
        check_image_axis(n, TZ, s, e)
    assert np.all(cover == 1), (n, cover)


def check_image_axis(n, T, s, e):
    """Imaging kernel: halo 2 window from window_start(s, 2), D1 taps of every cell."""
    wsi = window_start(s, H_IMG)
    Wi = T + 2 * H_IMG
    loaded_i = set(p for p in range(wsi, wsi + Wi) if 0 <= p < n)
    for g in range(s, e):
        for o in d1_offsets(g, n):
            assert 0 <= g + o - wsi < Wi and (g + o) in loaded_i, ("image", n, T, s, g, o)


@pytest.mark.parametrize("T", [16, 32])
def test_tiles_partition_and_windows(T):
    for n in list(range(12, 3 * T + 20)) + [201, 300, 500, 601, 1000, 1700, 3000]:
        check_x_axis(n, T)


def test_quad_tiles_along_z():
    for n in list(range(12, 300)) + [501, 601, 1000, 3001]:
        check_z_axis(n)


def xcd_remap(bid, nb):
    q, r = nb >> 3, nb & 7
    xcd, local = bid & 7, bid >> 3
    return xcd * (q + 1) + local if xcd < r else r * (q + 1) + (xcd - r) * q + local


def test_xcd_remap_is_a_permutation():
    for nb in list(range(1, 300)) + [4096, 27001]:
        m = [xcd_remap(b, nb) for b in range(nb)]
        assert sorted(m) == list(range(nb))
