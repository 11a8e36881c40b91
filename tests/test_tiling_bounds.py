"""CPU emulation of the step and imaging kernels' index arithmetic.

Mirrors the box decomposition of the engine (fwi2d_amd/csrc/engine.hip: build_boxes)
and the index arithmetic of k_step / k_image (fwi2d_amd/csrc/kernels.hip), and checks,
for many grid shapes and every box, that
  * the boxes partition the grid and fit their class's shape;
  * core boxes (compiled without closure code) only ever take the interior paths;
  * each stencil read (SBP interior and closure rows, first-derivative products, SAT
    rows, extraction footprints) lands inside the LDS arrays and on a grid cell the box
    loaded, so no kernel reads outside its shared-memory windows or uses a zero-filled
    halo cell as data.
The 2-D stencils are separable, so each axis of a box is checked on its own.
"""
import numpy as np
import pytest

TZ, H_IMG = 64, 2
CORE_X0, CORE_Z0 = 6, 8
SHAPES = {"core16": (16, 64, False), "core32": (32, 64, False), "band_z": (64, 16, True),
          "band_z8": (120, 8, True), "band_x": (8, 128, True)}


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
    """product_range of kernels.hip: the products a box writing [s, e) uses."""
    lo, hi = max(0, s - 2), min(n, e + 2)
    if s < 4:
        lo, hi = 0, max(hi, 6)
    if e > n - 4:
        lo, hi = min(lo, n - 6), n
    return lo, hi


def tiles(n, T):
    nt = (n + T - 1) // T
    return [tile_range(k, nt, T, n) for k in range(nt)]


def pitch(n):
    """Device z pitch of the engine: n rounded up to 32 cells (128-byte columns)."""
    return (n + 31) // 32 * 32


def tile_range_quads(k, ntile, T, n):
    ldz = pitch(n)
    if ntile == 1:
        return 0, n
    if k == ntile - 1:
        return ldz - T, n
    s = k * T
    return s, min(s + T, ldz - T)


def tiles_quads(n, T):
    nt = (pitch(n) + T - 1) // T
    return [tile_range_quads(k, nt, T, n) for k in range(nt)]


def core_z_end(nz):
    return ((nz - 8) // 4) * 4


def build_boxes(nx, nz, tile_x):
    """Mirror of Engine::build_boxes: list of (class, sx, ex, sz, ez)."""
    out = []
    cx0, cx1 = CORE_X0, nx - CORE_X0
    cz0, cz1 = CORE_Z0, max(core_z_end(nz), CORE_Z0)

    def add(c, sx, ex, sz, ez):
        if ex > sx and ez > sz:
            out.append((c, sx, ex, sz, ez))
    core = "core16" if tile_x == 16 else "core32"
    for sx, ex in tiles(nx, tile_x):
        for sz, ez in tiles_quads(nz, TZ):
            add(core, max(sx, cx0), min(ex, cx1), max(sz, cz0), min(ez, cz1))
    for sz, ez in ((0, cz0), (cz1, nz)):
        thin = ez - sz <= 8
        for sx, ex in tiles(nx, 120 if thin else 64):
            add("band_z8" if thin else "band_z", sx, ex, sz, ez)
    for sz in range(cz0, cz1, 128):
        ez = min(sz + 128, cz1)
        add("band_x", 0, cx0, sz, ez)
        add("band_x", cx1, nx, sz, ez)
    return out


def check_x(sx, ex, TX, n, edge):
    assert 0 <= sx < ex <= n and ex - sx <= TX
    ws, W, I = window_start(sx, 4), TX + 8, TX + 4
    loaded = set(p for p in range(ws, ws + W) if 0 <= p < n)
    plo, phi = needed_products(sx, ex, n)
    need = set(range(plo, phi))
    assert phi - plo <= I, (n, sx, plo, phi)
    for p in need:
        fast = 4 <= p < n - 4
        assert fast or edge, ("closure product column in a core box", n, sx, p)
        for o in ([-2, -1, 1, 2] if fast else d1_offsets(p, n)):
            assert 0 <= p + o - ws < W and (p + o) in loaded, (n, sx, p, o)
    for g in range(sx, ex):
        fast = 6 <= g < n - 6
        assert fast or edge, ("closure column in a core box", n, sx, g)
        for o in ([-2, -1, 0, 1, 2] if fast else d2_offsets(g, n)):
            assert 0 <= g + o - ws < W and (g + o) in loaded, (n, sx, g, o)
        for o in ([-2, -1, 1, 2] if fast else d1_offsets(g, n)):
            assert 0 <= g + o - plo < I and (g + o) in need, (n, sx, g, o)
            assert (g + o) in loaded
        if g == 0:
            assert all(k in loaded for k in range(6))
        if g == n - 1:
            assert all(n - 1 - k in loaded for k in range(6))
        if 1 <= g <= n - 2:  # extraction footprint of a pivot owned by this box
            assert all((g + o) in loaded and 0 <= g + o - ws < W for o in (-1, 0, 1))


def z_windows(sz, TZb):
    """(state window start, rows, product start, rows) of a box (k_step's TileInfo)."""
    if TZb == 8:  # compact windows of the 8-row bands
        return (0, 16, 0, 12) if sz == 0 else (sz - 8, 16, sz - 4, 12)
    return sz - 8, TZb + 16, sz - 4, TZb + 8


def check_z(sz, ez, TZb, n, edge):
    ldz = pitch(n)
    assert sz % 4 == 0 and 0 <= sz < ez <= n and ez - sz <= TZb
    ws, WZ, pz0, PZ = z_windows(sz, TZb)
    loaded = set(p for p in range(ws, ws + WZ) if 0 <= p < n)  # padding rows are not data
    plo, phi = needed_products(sz, ez, n)
    need = set(range(plo, phi))
    assert plo >= pz0 and phi <= pz0 + PZ, (n, sz, plo, phi)
    for lq in range(PZ // 4):
        z0 = pz0 + 4 * lq
        if not (z0 + 3 >= plo and z0 < phi):
            continue
        assert 0 <= z0 and z0 + 3 < ldz  # coefficient / snapshot float4 reads
        fast = z0 >= 4 and z0 + 3 < n - 4
        assert fast or edge, ("closure product quad in a core box", n, sz, z0)
        if fast:
            assert ws <= z0 - 4 and z0 + 7 < ws + WZ  # 12-row LDS read
            for k in range(4):
                if z0 + k in need:
                    assert all(z0 + k + o in loaded for o in (-2, -1, 1, 2))
        else:
            for k in range(4):
                g = z0 + k
                if 0 <= g < n:
                    for o in d1_offsets(g, n):
                        assert 0 <= g + o - ws < WZ and (g + o) in loaded, (n, sz, g, o)
    for j in range(TZb // 4):
        z0 = sz + 4 * j
        if z0 >= ez:
            continue
        assert z0 + 3 < ldz
        fast = z0 >= 6 and z0 + 3 < n - 6
        assert fast or edge, ("closure quad in a core box", n, sz, z0)
        if fast:
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
                    assert 0 <= g + o - ws < WZ and (g + o) in loaded, (n, sz, g, o)
                for o in d1_offsets(g, n):
                    assert 0 <= g + o - pz0 < PZ and (g + o) in need, (n, sz, g, o)
    for g in range(sz, ez):
        if g == 0:
            assert all(k in loaded for k in range(6))
        if g == n - 1:
            assert all(n - 1 - k in loaded for k in range(6))
        if 1 <= g <= n - 2:
            assert all((g + o) in loaded and 0 <= g + o - ws < WZ for o in (-1, 0, 1))


def check_grid(nx, nz, tile_x):
    cover = np.zeros((nx, nz), int)
    for c, sx, ex, sz, ez in build_boxes(nx, nz, tile_x):
        TX, TZb, edge = SHAPES[c]
        cover[sx:ex, sz:ez] += 1
        check_x(sx, ex, TX, nx, edge)
        check_z(sz, ez, TZb, nz, edge)
    assert np.all(cover == 1), (nx, nz)


SIZES = list(range(12, 90)) + [129, 131, 201, 300, 500, 601, 1000, 1700, 3000]


@pytest.mark.parametrize("tile_x", [16, 32])
def test_boxes_partition_and_windows_x(tile_x):
    for nx in SIZES:
        check_grid(nx, 70, tile_x)


def test_boxes_partition_and_windows_z():
    for nz in list(range(12, 300)) + [501, 601, 1000, 3001]:
        check_grid(40, nz, 16)


def check_image_axis(n, T, quads):
    """Imaging kernel (used without the imaging lag): halo 2 windows on the tile grid."""
    for s, e in (tiles_quads(n, T) if quads else tiles(n, T)):
        wsi = window_start(s, H_IMG)
        Wi = T + 2 * H_IMG
        loaded_i = set(p for p in range(wsi, wsi + Wi) if 0 <= p < n)
        for g in range(s, e):
            for o in d1_offsets(g, n):
                assert 0 <= g + o - wsi < Wi and (g + o) in loaded_i, ("image", n, T, s, g, o)


def check_img_tiles(nx, nz, T):
    """k_img: TX x 64 tiles of the tile grid, windows with halo 2 in x and 4 in z (quad
    reads of rows z0-4 .. z0+7); a tile with every cell 4 or more from the edges takes
    the interior stencils only."""
    ldz = pitch(nz)
    for sx, ex in tiles(nx, T):
        wsx, WX = window_start(sx, 2), T + 4
        lx = set(p for p in range(wsx, wsx + WX) if 0 <= p < nx)
        for sz, ez in tiles_quads(nz, 64):
            wsz, WZ = sz - 4, 72
            lz = set(p for p in range(wsz, wsz + WZ) if 0 <= p < nz)
            inner = sx >= 4 and ex <= nx - 4 and sz >= 4 and ez <= nz - 4
            for g in range(sx, ex):
                fast = 4 <= g < nx - 4
                assert fast or not inner
                for o in ([-2, -1, 1, 2] if fast else d1_offsets(g, nx)):
                    assert 0 <= g + o - wsx < WX and (g + o) in lx, ("img x", nx, T, sx, g, o)
            for z0 in range(sz, ez, 4):
                assert z0 + 3 < ldz
                fast = z0 >= 4 and z0 + 3 < nz - 4
                assert fast or not inner
                if fast:
                    assert wsz <= z0 - 4 and z0 + 7 < wsz + WZ
                    assert all(z0 + k + o in lz for k in range(4) for o in (-2, -1, 1, 2))
                else:
                    for k in range(4):
                        g = z0 + k
                        if g < nz:
                            for o in d1_offsets(g, nz):
                                assert 0 <= g + o - wsz < WZ and (g + o) in lz, ("img z", nz, sz, g, o)


@pytest.mark.parametrize("T", [16, 32])
def test_img_tiles(T):
    for nx in SIZES:
        check_img_tiles(nx, 70, T)
    for nz in list(range(12, 300)) + [501, 1000]:
        check_img_tiles(40, nz, T)


def test_image_tiles():
    for n in SIZES:
        check_image_axis(n, 16, False)
        check_image_axis(n, 32, False)
    for n in list(range(12, 300)) + [501, 1000]:
        check_image_axis(n, TZ, True)


THREADS, WAVE = 256, 64


def shape_consts(TX, TZb):
    """StepShape of kernels.hip: lane mapping constants of one box class."""
    rowmaj = TZb <= 16
    WQ, PQ, IX = (TZb + 16) // 4, (TZb + 8) // 4, TX + 4
    if TZb == 8:
        WQ, PQ = 4, 3
    WZ = 4 * ((WQ | 1) if rowmaj else WQ)
    PZ = 4 * ((PQ | 1) if rowmaj else PQ)
    IXP = (IX + 63) // 64 * 64 if rowmaj else IX
    BX = (TX + 63) // 64 * 64 if rowmaj else TX
    QC = TZb // 4
    return dict(rowmaj=rowmaj, WQ=WQ, PQ=PQ, IX=IX, WZ=WZ, PZ=PZ, IXP=IXP, QC=QC, WX=TX + 8, BX=BX,
                NQ2=(IXP * PQ + THREADS - 1) // THREADS, NQ3=(BX * QC + THREADS - 1) // THREADS)


def out_quad(c, TX, i):
    """(column, quad row) of new-level slot i (out_quad of kernels.hip); a column at or
    past the box width is an idle lane."""
    if c["rowmaj"]:
        return i % c["BX"], i // c["BX"]
    return i // c["QC"], i % c["QC"]


def prod_slot(c, i):
    """(column, quad row, valid) of product slot i (prod_slot of kernels.hip)."""
    if c["rowmaj"]:
        lq, lc = i // c["IXP"], i % c["IXP"]
        return lc, lq, lc < c["IX"] and lq < c["PQ"]
    return i // c["PQ"], i % c["PQ"], i < c["IX"] * c["PQ"]


@pytest.mark.parametrize("cls", sorted(SHAPES))
def test_lane_mappings(cls):
    """Every new-level quad and product quad of a box has exactly one slot; in row-major
    boxes the quad row of a slot (made a scalar with readfirstlane in the kernel) is the
    same on all 64 lanes of its wavefront; LDS column strides hold the windows."""
    TX, TZb, _ = SHAPES[cls]
    c = shape_consts(TX, TZb)
    assert c["rowmaj"] or (TX * c["QC"]) % THREADS == 0
    seen = set()
    for i in range(c["NQ3"] * THREADS):
        col, row = out_quad(c, TX, i)
        assert 0 <= col < c["BX"] and 0 <= row < c["QC"]
        if col >= TX:
            continue
        assert (col, row) not in seen
        seen.add((col, row))
        if c["rowmaj"]:
            assert out_quad(c, TX, i - i % WAVE)[1] == row
    assert len(seen) == TX * c["QC"]
    seen = set()
    for i in range(c["NQ2"] * THREADS):
        lc, lq, ok = prod_slot(c, i)
        if c["rowmaj"]:
            assert prod_slot(c, i - i % WAVE)[1] == lq
        if ok:
            assert (lc, lq) not in seen
            seen.add((lc, lq))
    assert len(seen) == c["IX"] * c["PQ"]
    _, wrows, _, prows = z_windows(8, TZb)
    assert c["WZ"] >= wrows and c["PZ"] >= prows and c["WZ"] % 4 == 0 and c["PZ"] % 4 == 0
    if c["rowmaj"]:
        # column-strided 16-byte LDS reads of consecutive lanes: odd quad strides
        assert (c["WZ"] // 4) % 2 == 1 and (c["PZ"] // 4) % 2 == 1


def xcd_remap(bid, nb):
    q, r = nb >> 3, nb & 7
    xcd, local = bid & 7, bid >> 3
    return xcd * (q + 1) + local if xcd < r else r * (q + 1) + (xcd - r) * q + local


def test_xcd_remap_is_a_permutation():
    for nb in list(range(1, 300)) + [4096, 27001]:
        m = [xcd_remap(b, nb) for b in range(nb)]
        assert sorted(m) == list(range(nb))
