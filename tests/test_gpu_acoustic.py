"""The acoustic engine (fwi2d_acoustic_*, nl_we_op_a::propagate, src/cpp/we_op.cpp:1937-2135)
against the FP64 oracle: receiver traces within the reference algorithm's own FP32 gap
(at least 1e-5) and raw (K, 1/rho) gradients within 1e-4, for every boundary kind (free
surface, absorbing, rigid), with and without tapers, ragged grids and sub > 1; the batch
of shots equals each shot alone bit for bit."""
import numpy as np
import pytest

from fwi2d_amd import workloads as W
from oracle.oracle import grid_struct

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def lib():
    import fwi2d_amd
    fwi2d_amd.load()
    return fwi2d_amd


def case(nx=90, nz=70, nt=400, bc=(1, 2, 2, 2), taper=10, ns=2, sub=1, seed=0):
    h, dt = 0.01, 0.001
    rng = np.random.default_rng(seed)
    x = np.arange(nx)[:, None] * h
    z = np.arange(nz)[None, :] * h
    vp = 1.8 + 1.2 * z + 0.2 * np.exp(-((x - 0.4) ** 2 + (z - 0.3) ** 2) / 0.01) + 0 * x
    rho = 2.0 + 0.2 * np.sin(2 * np.pi * x / 0.6) * np.cos(2 * np.pi * z / 0.5)
    model = np.stack([rho * vp ** 2, 1 / rho])
    sxz = np.stack([rng.uniform(0.1, (nx - 10) * h, ns), rng.uniform(0.02, 0.15, ns)], 1)
    rxz = [np.stack([np.linspace(0.03, (nx - 4) * h, 12 + 3 * s), np.full(12 + 3 * s, 0.02 + 0.01 * s)], 1)
           for s in range(ns)]
    grid = dict(nx=nx, nz=nz, dx=h, dz=h, nt=nt, dt=dt, sub=sub, bc_top=bc[0], bc_bottom=bc[1], bc_left=bc[2],
                bc_right=bc[3], taper_top=0 if bc[0] != 2 else taper, taper_bottom=taper, taper_left=taper,
                taper_right=taper, taper_strength=0.05, vti=0, gdamp_tmin=0.0, gdamp_tmax=0.0, gdamp_power=1.0)
    src = np.stack([W.ricker(nt, dt, 15.0 + 3 * s) for s in range(ns)])
    return grid, model, sxz, rxz, src


def oracle_run(orc, grid, model, sxz, rxz, src, adj=None, stiffness=1.0):
    g = grid_struct(grid)
    nx, nz, h = grid["nx"], grid["nz"], grid["dx"]
    alpha = 0.2508560249 / max(stiffness, 1.0)
    nt, sub = grid["nt"], grid["sub"]
    outs, grads = [], np.zeros((2, nx, nz))
    r0 = 0
    for s in range(len(sxz)):
        inj = orc.delta_m3(nx, nz, h, h, 0, 0, [sxz[s, 0]], [sxz[s, 1]])
        ext = orc.delta_m3(nx, nz, h, h, 0, 0, rxz[s][:, 0], rxz[s][:, 1])
        full = np.zeros((1 + nt // sub, nx, nz)) if sub > 0 else None
        outs.append(orc.propagate_acoustic(g, model, inj, src[s:s + 1], ext=ext, full=full, alpha=alpha))
        if adj is not None:
            n = len(rxz[s])
            orc.propagate_acoustic(g, model, ext, adj[r0:r0 + n], full=full, grad=grads, alpha=alpha)
            r0 += n
    return np.concatenate(outs, 0), grads


def engine(lib, grid, model, sxz, rxz, stiffness=1.0):
    from fwi2d_amd.lib import Fwi2dGrid
    g = Fwi2dGrid()
    for k, v in grid.items():
        setattr(g, k, v)
    e = lib.AcousticEngine(g, stiffness)
    e.set_model(model)
    e.set_shots(sxz, rxz)
    return e


@pytest.mark.parametrize("bc,taper,nx,nz,sub", [((1, 2, 2, 2), 10, 90, 70, 1), ((2, 2, 2, 2), 0, 61, 47, 1),
                                                  ((3, 3, 3, 3), 0, 37, 29, 1), ((1, 3, 2, 1), 8, 130, 83, 2)])
def test_acoustic_traces_and_gradient_match_oracle(lib, orc64, orc32, bc, taper, nx, nz, sub):
    grid, model, sxz, rxz, src = case(nx=nx, nz=nz, bc=bc, taper=taper, sub=sub)
    e = engine(lib, grid, model, sxz, rxz)
    d = e.forward(src, keep_snapshots=True)
    nr = sum(len(r) for r in rxz)
    adj = np.random.default_rng(nx).standard_normal((nr, grid["nt"])) * np.hanning(grid["nt"])
    ref, gref = oracle_run(orc64, grid, model, sxz, rxz, src, adj=adj)
    gap = rel(oracle_run(orc32, grid, model, sxz, rxz, src)[0], ref)
    err = rel(d, ref)
    tol = max(1e-5, 2 * gap)
    print("bc=%s taper=%d %dx%d sub=%d: traces rel err %.3e (tol %.1e)" % (bc, taper, nx, nz, sub, err, tol))
    assert err <= tol
    gk = e.adjoint(adj)
    for c, name in enumerate(("K", "1/rho")):
        errg = rel(gk[c], gref[c])
        print("  %s gradient rel err %.3e" % (name, errg))
        assert errg <= 1e-4
    e.close()


def test_acoustic_batch_equals_single_shots(lib):
    grid, model, sxz, rxz, src = case(ns=3)
    e = engine(lib, grid, model, sxz, rxz)
    d = e.forward(src)
    r0 = 0
    for s in range(3):
        one = engine(lib, grid, model, sxz[s:s + 1], rxz[s:s + 1])
        n = len(rxz[s])
        assert np.array_equal(one.forward(src[s:s + 1]), d[r0:r0 + n])
        r0 += n
        one.close()
    e.close()


def test_acoustic_stiffness_and_errors(lib, orc64):
    grid, model, sxz, rxz, src = case(nx=50, nz=40, nt=200)
    e = engine(lib, grid, model, sxz, rxz, stiffness=2.5)
    ref, _ = oracle_run(orc64, grid, model, sxz, rxz, src, stiffness=2.5)
    assert rel(e.forward(src), ref) < 1e-5
    with pytest.raises(lib.Fwi2dError, match="snapshots"):
        e.adjoint(np.zeros((sum(len(r) for r in rxz), grid["nt"])))
    e.close()
    bad = dict(grid, bc_top=4)
    with pytest.raises(lib.Fwi2dError, match="boundaries"):
        engine(lib, bad, model, sxz, rxz)


def op_case(seismotype=0, bc=(1, 2, 2, 2), dt_data=0.002, nt_data=201, nx=120, nz=80, stiffness=1.0):
    """nl_we_op_a at the operator level: model (vp, rho), sources [1][ns][nt_data],
    data [1][nr][nt_data]."""
    h = 0.01
    x = np.arange(nx)[:, None] * h
    z = np.arange(nz)[None, :] * h
    vp = 1.8 + 1.2 * z + 0.25 * np.exp(-((x - 0.6) ** 2 + (z - 0.45) ** 2) / 0.01)
    rho = 2.0 + 0.15 * np.sin(2 * np.pi * x / 0.55) * np.cos(2 * np.pi * z / 0.4)
    true = np.stack([vp, rho])
    start = np.stack([1.8 + 1.2 * z + 0 * x, np.full((nx, nz), 2.05)])
    rx = np.stack([np.arange(4, nx - 4, 3) * h, np.full(len(range(4, nx - 4, 3)), 0.04)], 1)
    rx2 = np.stack([np.full(15, 0.9), np.linspace(0.1, 0.7, 15)], 1)
    prob = W.base_problem(nx, nz, h, h, nt_data, dt_data, [[0.3, 0.1], [0.75, 0.06], [1.0, 0.3]], [rx, rx2, rx[::2]],
                          taper=12, bc=bc, seismotype=seismotype, dt=0.001, sub=-1)
    prob["acoustic"] = 1
    prob["free_surface_stiffness"] = stiffness
    src = np.stack([W.ricker(nt_data, dt_data, f) for f in (12.0, 9.0, 14.0)])[None]
    return prob, true.astype(np.float32), start.astype(np.float32), src.astype(np.float32)


@pytest.mark.parametrize("seismotype,bc,dt_data,stiffness,resampling", [(0, (1, 2, 2, 2), 0.002, 1.0, 0),
                                                                        (1, (2, 3, 2, 1), 0.001, 1.0, 0),
                                                                        (0, (1, 2, 2, 2), 0.003, 2.0, 0),
                                                                        (1, (1, 2, 2, 2), 0.0025, 1.05, 1)])
def test_acoustic_operator_matches_oracle(lib, orc64, seismotype, bc, dt_data, stiffness, resampling):
    """WeOp with prob["acoustic"] (nl_we_op_a): apply_forward (sinc or linear resampling,
    also at a non-integer ratio of the sampling rates, time derivative, dt/dt_data
    scaling) and apply_jacobianT (H quadrature, chain rule to vp and rho) against the
    oracle's operator within 1e-4."""
    prob, true, start, src = op_case(seismotype, bc, dt_data, stiffness=stiffness)
    prob["resampling"] = resampling
    dobs, _ = orc64.apply_forward(orc64.make_op(prob), true, src)
    d0, full = orc64.apply_forward(orc64.make_op(prob), start, src, keep_full=True)
    op = lib.WeOp(prob, src)
    assert op.np == 2 and op.ncd == 1
    d0k = op.apply_forward(start)
    assert d0k.shape == d0.shape == (1, op.nr, prob["mp"]["nt_data"])
    err_d = rel(d0k, d0)
    print("seismotype=%d bc=%s dt_data=%g: forward rel err %.3e" % (seismotype, bc, dt_data, err_d))
    assert err_d <= 1e-4
    rw = (d0 - dobs) * dt_data
    rw[..., 0] *= 0.5
    rw[..., -1] *= 0.5
    gref = orc64.apply_jacobianT(orc64.make_op(prob), start, rw, full)
    gk = op.apply_jacobianT(start, rw)
    assert gk.shape == (2,) + start.shape[1:]
    for c, name in enumerate(("vp", "rho")):
        err = rel(gk[c], gref[c])
        print("  %s gradient rel err %.3e" % (name, err))
        assert err <= 1e-4
    # the least-squares objective on the acoustic operator: 1/2 dt |r|^2 (trapezoid) of its
    # own residual, and the gradient J^T of that residual with the trapezoid weights (the
    # residual is a few percent of the data, so its FP32 error is compared on itself)
    nl = lib.Nlls(op, dobs)
    f, g = nl.evaluate(start)
    rk = np.asarray(d0k, np.float64) - dobs
    assert f == pytest.approx(orc64.nlls_functional(rk, dt_data), rel=1e-5)
    rwk = rk * dt_data
    rwk[..., 0] *= 0.5
    rwk[..., -1] *= 0.5
    grefk = orc64.apply_jacobianT(orc64.make_op(prob), start, rwk, full)
    for c in range(2):
        err = rel(g[c], grefk[c])
        print("  objective gradient %d rel err %.3e" % (c, err))
        assert err <= 1e-4
    nl.close()
    op.close()


def test_acoustic_shot_groups_and_interior_tiles(lib, orc64):
    """A grid with an interior rectangle of tiles and 65 shots, which the step kernel runs
    3 to a workgroup (the last group holding 2): shots from the first, a middle and the
    last group equal the same shots run alone bit for bit, and one of them the oracle."""
    grid, model, sxz, rxz, src = case(nx=300, nz=260, nt=300, ns=65, taper=12, seed=4)
    e = engine(lib, grid, model, sxz, rxz)
    d = e.forward(src)
    offs = np.concatenate([[0], np.cumsum([len(r) for r in rxz])])
    for s in (0, 31, 64):
        one = engine(lib, grid, model, sxz[s:s + 1], rxz[s:s + 1])
        assert np.array_equal(one.forward(src[s:s + 1]), d[offs[s]:offs[s + 1]])
        one.close()
    ref, _ = oracle_run(orc64, grid, model, sxz[64:65], rxz[64:65], src[64:65])
    err = rel(d[offs[64]:offs[65]], ref)
    print("shot 64 of 65: traces rel err %.3e" % err)
    assert err <= 2e-5
    e.close()


@pytest.mark.parametrize("nt,sub", [(400, 1), (400, 2), (400, 3), (401, 4), (72, 2)])
def test_acoustic_checkpointed_snapshots_equal_the_whole_history(lib, nt, sub):
    """With the snapshot budget at 0 the forward run keeps checkpoints of segments of about
    sqrt(2 nsnap) snapshots and the adjoint run recomputes them; the gradient equals the
    one imaged from the whole history bit for bit, including nt a multiple of sub (the
    reference's zero last slot, at 72 steps and sub 2 alone in the last segment)."""
    grid, model, sxz, rxz, src = case(nx=70, nz=60, nt=nt, sub=sub, ns=3, seed=sub)
    nr = sum(len(r) for r in rxz)
    adj = np.random.default_rng(nt + sub).standard_normal((nr, nt)).astype(np.float32)
    grads, traces = [], []
    for budget in (1 << 20, 0):
        e = engine(lib, grid, model, sxz, rxz)
        e.set_option("snapshot_budget", budget)
        traces.append(e.forward(src, keep_snapshots=True))
        grads.append(e.adjoint(adj))
        assert e.stat("whole_snapshots") == (1 if budget else 0)
        n = 1 + nt // sub
        K = e.stat("segment")
        if budget == 0:
            # every segment but the last is recomputed once, from one snapshot before its
            # own to one after: about one forward run
            assert 0 < e.stat("recompute_steps") <= nt + 2 * sub * e.stat("nseg")
            assert e.stat("nseg") == (n + K - 1) // K and K * K >= 2 * n
        e.close()
    assert np.array_equal(traces[0], traces[1])
    assert np.array_equal(grads[0], grads[1])
    assert np.abs(grads[0]).max() > 0


def test_acoustic_full_size_properties(lib):
    """Size-independent properties of the acoustic engine on the 64-shot 500 x 1700 grid
    of the batched benchmark (a short time axis): scaling in the source (powers of two
    scale FP32 exactly away from the subnormal range: 1e-7), superposition of two source
    sets, one shot alone equal to the batch bit for bit, scaling of the gradient in the
    adjoint source, and the checkpointed gradient equal to the whole-history one."""
    nz, nx, h, dt, nt, ns = 500, 1700, 0.01, 0.001, 240, 64
    m = W.gradient_model(nx, nz, h, h, seed=0)
    model = np.stack([m[2] * m[0] ** 2, 1 / m[2]]).astype(np.float32)
    sxz, rxz, freqs = W.config2_geometry(ns, nx, h)
    prob = W.base_problem(nx, nz, h, h, nt, dt, sxz, rxz, dt=dt, sub=1)
    wav = np.stack([W.ricker(nt, dt, f) for f in freqs]).astype(np.float32)
    s2 = (-0.5 * wav[:, ::-1]).copy()
    e = lib.AcousticEngine(lib.grid_from(prob, nt, dt, 1))
    e.set_model(model)
    e.set_shots(sxz, rxz)
    assert e.stat("shots_per_block") == 8  # the benchmark's shot groups
    d1 = e.forward(wav)
    err = rel(e.forward(2 * wav), 2 * d1.astype(np.float64))
    print("scaling rel err %.2e" % err)
    assert err < 1e-7
    d2 = e.forward(s2)
    err = rel(e.forward(wav + s2), d1.astype(np.float64) + d2)
    print("superposition rel err %.2e" % err)
    assert err < 1e-5
    k = 37
    one = lib.AcousticEngine(lib.grid_from(prob, nt, dt, 1))
    one.set_model(model)
    one.set_shots(sxz[k:k + 1], rxz[k:k + 1])
    r0 = sum(len(r) for r in rxz[:k])
    assert np.array_equal(one.forward(wav[k:k + 1]), d1[r0:r0 + len(rxz[k])])
    one.close()
    adj = (np.random.default_rng(4).standard_normal(d1.shape) * np.hanning(nt)).astype(np.float32)
    e.forward(wav, keep_snapshots=True)
    g1 = e.adjoint(adj)
    err = rel(e.adjoint(2 * adj), 2 * g1.astype(np.float64))
    print("gradient scaling rel err %.2e" % err)
    assert err < 1e-7
    assert np.all(np.isfinite(g1)) and np.linalg.norm(g1) > 0
    e.set_option("snapshot_budget", 0)
    e.forward(wav, keep_snapshots=True)
    assert e.stat("whole_snapshots") == 0
    assert np.array_equal(e.adjoint(adj), g1)
    e.close()


def test_acoustic_shot_without_receivers_and_small_grids(lib, orc64):
    """A shot with no receivers (a ragged acquisition) leaves the other shots' traces and
    the gradient as the acquisition without it gives them; grids below 12 cells fail."""
    grid, model, sxz, rxz, src = case(nx=80, nz=60, nt=300, ns=3, seed=7)
    rxz0 = [rxz[0], np.zeros((0, 2)), rxz[2]]
    e = engine(lib, grid, model, sxz, rxz0)
    d = e.forward(src, keep_snapshots=True)
    assert d.shape == (len(rxz[0]) + len(rxz[2]), 300)
    ref = engine(lib, grid, model, sxz[[0, 2]], [rxz[0], rxz[2]])
    assert np.array_equal(d, ref.forward(src[[0, 2]], keep_snapshots=True))
    adj = np.random.default_rng(2).standard_normal(d.shape).astype(np.float32)
    assert np.array_equal(e.adjoint(adj), ref.adjoint(adj))
    e.close()
    ref.close()
    with pytest.raises(lib.Fwi2dError, match="nx, nz >= 12"):
        engine(lib, dict(grid, nz=11), model[:, :, :11], sxz, rxz)


@pytest.mark.parametrize("nx,nz,nt,sub,ns", [(90, 70, 300, 1, 3), (300, 260, 200, 2, 65), (61, 47, 250, 3, 2)])
def test_acoustic_fused_imaging_equals_separate(lib, nx, nz, nt, sub, ns):
    """The imaging inside the adjoint step kernel (option image_fused) and the separate
    imaging kernel (the default) sum the same terms in the same order per shot group:
    the gradients agree to rounding (measured: bit for bit), for whole and checkpointed
    snapshot histories."""
    grid, model, sxz, rxz, src = case(nx=nx, nz=nz, nt=nt, sub=sub, ns=ns, seed=nx)
    nr = sum(len(r) for r in rxz)
    adj = (np.random.default_rng(3).standard_normal((nr, nt)) * np.hanning(nt)).astype(np.float32)
    for budget in (1 << 20, 0):
        grads = []
        for fused in (1, 0):
            e = engine(lib, grid, model, sxz, rxz)
            e.set_option("snapshot_budget", budget)
            e.set_option("image_fused", fused)
            assert e.stat("image_fused") == fused
            e.forward(src, keep_snapshots=True)
            grads.append(e.adjoint(adj))
            e.close()
        err = max(rel(grads[0][c], grads[1][c]) for c in range(2))
        print("%dx%d sub=%d ns=%d budget=%d: fused vs separate imaging rel err %.2e" % (nx, nz, sub, ns, budget, err))
        assert err <= 1e-6


@pytest.mark.parametrize("fused", [0, 1])
def test_acoustic_gradient_time_window(lib, orc64, fused):
    """The imaging's time window (par.gdamp_tmin/tmax/power, we_op.cpp:1903-1906): the
    snapshots between tmin and tmax weighted by cos^2 of the window's phase; the raw
    gradients against the FP64 oracle within 1e-4, with the imaging separate and fused,
    and the window changes them by far more than that."""
    grid, model, sxz, rxz, src = case(nx=90, nz=70, nt=400, sub=2, seed=5)
    nr = sum(len(r) for r in rxz)
    adj = (np.random.default_rng(9).standard_normal((nr, 400)) * np.hanning(400)).astype(np.float32)
    gw = dict(grid, gdamp_tmin=0.05, gdamp_tmax=0.3, gdamp_power=1.0)
    _, gref = oracle_run(orc64, gw, model, sxz, rxz, src, adj=adj)
    _, g0 = oracle_run(orc64, grid, model, sxz, rxz, src, adj=adj)
    e = engine(lib, gw, model, sxz, rxz)
    e.set_option("image_fused", fused)
    e.forward(src, keep_snapshots=True)
    gk = e.adjoint(adj)
    e.close()
    for c in range(2):
        err, change = rel(gk[c], gref[c]), rel(g0[c], gref[c])
        print("fused=%d component %d: rel err %.2e, window changes the gradient by %.2e" % (fused, c, err, change))
        assert err <= 1e-4 and change > 1e-2


def test_acoustic_origin_and_add(lib, orc64):
    """A grid whose origin is not zero (the sources and receivers shifted with it) gives
    the traces of the unshifted problem; at the operator level add = 1 adds F(m) and
    J^T d to what the output holds (the mathematical add of the C ABI)."""
    grid, model, sxz, rxz, src = case(nx=80, nz=60, nt=300, ns=2, seed=11)
    ox, oz = 1.25, 0.4
    e0 = engine(lib, grid, model, sxz, rxz)
    e1 = lib.AcousticEngine(lib.grid_from(grid, grid["nt"], grid["dt"], grid["sub"]), 1.0)
    e1.set_model(model)
    e1.set_shots(sxz + [ox, oz], [r + [ox, oz] for r in rxz], ox, oz)
    assert np.array_equal(e0.forward(src), e1.forward(src))
    e0.close()
    e1.close()
    prob, true, start, srcs = op_case(0, (1, 2, 2, 2), 0.002)
    op = lib.WeOp(prob, srcs)
    d = op.apply_forward(start)
    base = np.random.default_rng(1).standard_normal(d.shape).astype(np.float32)
    assert np.allclose(op.apply_forward(start, add_to=base), base + d, rtol=0, atol=1e-6 * np.abs(d).max())
    g = op.apply_jacobianT(start, d)
    gb = np.random.default_rng(2).standard_normal(g.shape).astype(np.float32) * np.abs(g).max()
    assert np.allclose(op.apply_jacobianT(start, d, add_to=gb), gb + g, rtol=0, atol=1e-6 * np.abs(g).max())
    op.close()
