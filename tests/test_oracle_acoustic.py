"""Pins of the oracle's acoustic operator (nl_we_op_a, src/cpp/we_op.cpp:1863-2135; the
acoustic SAT terms and their boundary scaling, src/cpp/spatial_operators.cpp:551-860), on
the CPU.

  * asat_dirichlet/absorbing/neumann on every side and asat_scale_boundaries for every
    combination of absorbing sides equal the reference's own functions compiled from
    source (oracle/_ref) to rounding.
  * The spatial operator D2x(1/rho) + D2z(1/rho) + SAT with free-surface (Dirichlet) or
    rigid (Neumann) sides is symmetric and negative semi-definite in the H inner product,
    so the leapfrog conserves the discrete energy (H/K weights) to rounding.
  * In a homogeneous medium the propagation (injection, SBP operator, absorbing SAT,
    taper, extraction) matches the closed-form 2-D Green's function.
  * The gradients (K with its -1/K^2 factor, 1/rho) against finite differences of the
    L2 misfit: a sign, a missing dt or factor, or a shifted snapshot fails them.
"""
import numpy as np
import pytest

from fwi2d_amd import workloads as W
from oracle.oracle import grid_struct
from tests.analytic import acoustic_greens_traces


def hvec(n, h=1.0):
    return h * np.array([17 / 48, 59 / 48, 43 / 48, 49 / 48] + [1] * (n - 8) + [49 / 48, 43 / 48, 59 / 48, 17 / 48])


def ac_model(nx, nz, seed=3):
    rng = np.random.default_rng(seed)
    vp = 2.0 + 0.3 * rng.random((nx, nz))
    rho = 1.8 + 0.4 * rng.random((nx, nz))
    return np.stack([rho * vp ** 2, 1 / rho])


@pytest.mark.parametrize("side", [0, 1, 2, 3])
@pytest.mark.parametrize("kind", [1, 2, 3])
def test_asat_matches_reference(orc64, refsbp, side, kind):
    nx, nz, dx, dz, dt = 17, 13, 0.011, 0.009, 0.001
    K, ri = ac_model(nx, nz)
    rng = np.random.default_rng(side * 7 + kind)
    u, up, out = rng.standard_normal((3, nx, nz))
    a = 0.2508560249 / 1.7
    got = orc64.asat(side, kind, u, up, out, K, ri, dx, dz, dt, a)
    ref = refsbp.asat(side, kind, u, up, out, K, ri, dx, dz, dt, a)
    assert np.abs(got - out).max() > 0  # the SAT did something
    assert np.allclose(got, ref, rtol=0, atol=1e-13 * np.abs(ref).max())


@pytest.mark.parametrize("flags", [tuple((m >> k) & 1 for k in range(4)) for m in range(16)])
def test_asat_scale_matches_reference(orc64, refsbp, flags):
    nx, nz = 15, 12
    K, ri = ac_model(nx, nz)
    u = np.random.default_rng(sum(flags)).standard_normal((nx, nz))
    got = orc64.asat_scale(u, K, ri, 0.01, 0.012, 0.001, *flags)
    ref = refsbp.asat_scale(u, K, ri, 0.01, 0.012, 0.001, *flags)
    assert np.allclose(got, ref, rtol=1e-14, atol=0)


def ac_rhs(orc, u, K, ri, dx, dz, kinds, a):
    out = orc.Dxx_var(u, ri, dx) + orc.Dzz_var(u, ri, dz)
    for side, kind in enumerate(kinds):
        out = orc.asat(side, kind, u, np.zeros_like(u), out, K, ri, dx, dz, 0.001, a)
    return out


A_FREE = 0.2508560249  # the Dirichlet SAT parameter at free_surface_stiffness 1 (we_op.cpp:1962)


def operator_matrix(orc, K, ri, dx, dz, kinds, a):
    nx, nz = K.shape
    N = nx * nz
    A = np.zeros((N, N))
    for k in range(N):
        e = np.zeros(N)
        e[k] = 1
        A[:, k] = ac_rhs(orc, e.reshape(nx, nz), K, ri, dx, dz, kinds, a).ravel()
    return np.outer(hvec(nx, dx), hvec(nz, dz)).ravel()[:, None] * A


@pytest.mark.parametrize("kinds", [(1, 1, 1, 1), (3, 3, 3, 3), (1, 3, 3, 1), (2, 2, 2, 2)])
def test_acoustic_operator_self_adjoint(orc64, kinds):
    """H (D2x(1/rho) + D2z(1/rho) + SAT) is symmetric and negative semi-definite (the
    absorbing SAT's spatial part, with the previous level zero, is the Neumann term; the
    free surface at stiffness 2: with cell-to-cell jumps of the coefficients the stability
    limit lies slightly above 1)."""
    M = operator_matrix(orc64, *ac_model(14, 13), 0.011, 0.009, kinds, A_FREE / 2)
    scale = np.abs(M).max()
    assert np.abs(M - M.T).max() < 1e-13 * scale
    assert np.linalg.eigvalsh((M + M.T) / 2).max() < 1e-13 * scale


def test_free_surface_penalty_is_the_stability_limit(orc64):
    """The reference's Dirichlet parameter 0.2508560249 / stiffness (stiffness >= 1 for the
    acoustic operator, we_op.cpp:16) is the exact stability limit of the SAT with constant
    coefficients: at stiffness 1 the largest eigenvalue of H A is zero to 1e-9 of the
    operator's scale, at 1.01 it is clearly negative, at 0.99 positive.  A wrong H or
    one-sided coefficient in the SAT moves the limit."""
    nx, nz = 14, 13
    K, ri = np.full((nx, nz), 8.0), np.full((nx, nz), 0.5)
    top = {}
    for stiff in (0.99, 1.0, 1.01):
        M = operator_matrix(orc64, K, ri, 0.011, 0.009, (1, 1, 1, 1), A_FREE / stiff)
        top[stiff] = np.linalg.eigvalsh((M + M.T) / 2).max() / np.abs(M).max()
    print(top)
    assert abs(top[1.0]) < 1e-9 and top[1.01] < -1e-3 and top[0.99] > 1e-3


def ac_problem(nx, nz, h, nt, dt, bc, taper, sub=1):
    return grid_struct(dict(nx=nx, nz=nz, dx=h, dz=h, nt=nt, dt=dt, sub=sub, bc_top=bc[0], bc_bottom=bc[1],
                            bc_left=bc[2], bc_right=bc[3], taper_top=taper, taper_bottom=taper, taper_left=taper,
                            taper_right=taper, taper_strength=0.05, vti=0, gdamp_tmin=0.0, gdamp_tmax=0.0,
                            gdamp_power=1.0))


@pytest.mark.parametrize("bc", [(1, 1, 1, 1), (3, 3, 3, 3), (1, 3, 3, 3)])
def test_acoustic_energy_conservation(orc64, bc):
    nx, nz, h, dt, nt = 24, 20, 0.01, 0.001, 400
    K, ri = ac_model(nx, nz)
    g = ac_problem(nx, nz, h, nt, dt, bc, 0)
    xi, zi, xw, zw = orc64.delta_m3(nx, nz, h, h, 0, 0, [0.113], [0.087])
    w = W.ricker(nt, dt, 25.0)
    w[80:] = 0.0
    full = np.zeros((1 + nt, nx, nz))
    orc64.propagate_acoustic(g, np.stack([K, ri]), (xi, zi, xw, zw), w[None], full=full, alpha=A_FREE / 2)
    wq = np.outer(hvec(nx, h), hvec(nz, h))
    E = []
    for n in range(100, nt - 1):
        u0, u1 = full[n], full[n + 1]
        kin = np.sum(wq / K * (u1 - u0) ** 2) / dt ** 2
        pot = -np.sum(wq * u1 * ac_rhs(orc64, u0, K, ri, h, h, bc, A_FREE / 2))
        E.append(kin + pot)
    E = np.array(E)
    assert E.min() > 0
    assert (E.max() - E.min()) / E.mean() < 1e-10


def test_acoustic_point_source_matches_greens_function(orc64):
    """A 4 Hz Ricker source at the centre of a 3 km square with absorbing sides, receivers
    0.25 km away (two off the grid), before any boundary arrival."""
    n, h, nt, dt, f0 = 301, 0.01, 1001, 0.001, 4.0
    v, rho = 2.0, 2.0
    model = np.stack([np.full((n, n), rho * v * v), np.full((n, n), 1 / rho)])
    g = ac_problem(n, n, h, nt, dt, (2, 2, 2, 2), 20, sub=0)
    xs = zs = (n - 1) * h / 2
    offsets = [(0.25, 0.0), (0.0, 0.25), (0.25 / np.sqrt(2), 0.25 / np.sqrt(2)), (-0.2, -0.15)]
    rxz = np.array([[xs + a, zs + b] for a, b in offsets])
    t = np.arange(nt) * dt
    a = (np.pi * f0 * (t - 1.2 / f0)) ** 2
    wav = (1.0 - 2.0 * a) * np.exp(-a)
    inj = orc64.delta_m3(n, n, h, h, 0, 0, [xs], [zs])
    ext = orc64.delta_m3(n, n, h, h, 0, 0, rxz[:, 0], rxz[:, 1])
    out = orc64.propagate_acoustic(g, model, inj, wav[None], ext=ext)
    ref = acoustic_greens_traces(wav, dt, v, rho, offsets)
    for r in range(len(offsets)):
        err = np.linalg.norm(out[r] - ref[r]) / np.linalg.norm(ref[r])
        print("receiver %d: rel err %.2e" % (r, err))
        assert err < 2e-3


def _acoustic_fd_errors(orc, h):
    """Relative errors (gradient - FD) / FD of dJ/dK and dJ/d(1/rho) of
    J = 1/2 sum_t h_t |d - d_obs|^2 (trapezoid weights h_t) on a 0.5 x 0.4 km problem with
    cell size h: the raw gradients with the H quadrature against central differences
    along smooth perturbations; the adjoint sources are the time-reversed, sign-flipped
    residuals (the reference's apply_jacobianT at equal sampling)."""
    from scipy.ndimage import gaussian_filter
    nx, nz = int(round(0.49 / h)) + 1, int(round(0.39 / h)) + 1
    dt, T = 0.001, 0.45
    nt = int(round(T / dt)) + 1
    x = np.arange(nx)[:, None] * h
    z = np.arange(nz)[None, :] * h
    vp_true = 1.8 + 1.5 * z + 0.2 * np.exp(-((x - 0.3) ** 2 + (z - 0.25) ** 2) / 0.004)
    vp0 = 1.8 + 1.5 * z + 0.0 * x
    rho = 2.0 + 0.1 * np.sin(2 * np.pi * x / 0.5) * np.cos(2 * np.pi * z / 0.4)
    g = ac_problem(nx, nz, h, nt, dt, (1, 3, 3, 3), 0)
    inj = orc.delta_m3(nx, nz, h, h, 0, 0, [0.25], [0.2])
    rx = np.linspace(0.05, 0.44, 40)
    ext = orc.delta_m3(nx, nz, h, h, 0, 0, rx, np.full(len(rx), 0.12))
    src = W.ricker(nt, dt, 10.0)[None]
    wt = np.full(nt, dt)
    wt[0] = wt[-1] = 0.5 * dt

    def model(vp, r):
        return np.stack([r * vp ** 2, 1 / r])

    dobs = orc.propagate_acoustic(g, model(vp_true, rho), inj, src, ext=ext, alpha=A_FREE / 2)

    def J(m):
        d = orc.propagate_acoustic(g, m, inj, src, ext=ext, alpha=A_FREE / 2)
        return 0.5 * np.sum(wt * (d - dobs) ** 2)

    m0 = model(vp0, rho)
    full = np.zeros((nt, nx, nz))
    d0 = orc.propagate_acoustic(g, m0, inj, src, ext=ext, full=full, alpha=A_FREE / 2)
    adj = -(d0 - dobs)[:, ::-1].copy()
    gr = np.zeros((2, nx, nz))
    orc.propagate_acoustic(g, m0, ext, adj, full=full, grad=gr, alpha=A_FREE / 2)
    Hq = np.outer(hvec(nx, h), hvec(nz, h))
    s = 0.01 * np.sin(2 * np.pi * x / 0.25 + 0.3) * np.cos(2 * np.pi * z / 0.2 + 0.1)
    errs = {}
    for c, name in enumerate(("K", "1/rho")):
        dm = np.zeros_like(m0)
        dm[c] = m0[c] * s
        fd = (J(m0 + 0.1 * dm) - J(m0 - 0.1 * dm)) / 0.2
        an = np.sum(Hq * gr[c] * dm[c])
        errs[name] = (an - fd) / fd
        print("h=%g %s: analytic %.6e fd %.6e rel %.2e" % (h, name, an, fd, errs[name]))
    return errs


def test_acoustic_gradient_finite_difference(orc64):
    """The K gradient is the exact discrete gradient (the snapshot second difference is the
    forward leapfrog's own); the 1/rho gradient, the wide D1^T D1 product against the
    narrow D2(1/rho) stencil of the propagation, differs by an amount that falls at least
    fourfold when the cell size halves (the acoustic counterpart of the elastic mu term): a
    missing factor or sign would not shrink with h."""
    e1 = _acoustic_fd_errors(orc64, 0.01)
    e2 = _acoustic_fd_errors(orc64, 0.005)
    assert abs(e1["K"]) < 1e-4 and abs(e2["K"]) < 1e-4
    # measured 8.6e-3 at 10 m and 1.8e-4 at 5 m on this smooth perturbation
    assert abs(e1["1/rho"]) < 0.03
    assert abs(e2["1/rho"]) < abs(e1["1/rho"]) / 4


# ---------------------------------------------------------------------------
# Operator level: nl_we_op_a through apply_forward / apply_jacobianT
# (we_op.cpp:904-1239 with nmodels = 2)
# ---------------------------------------------------------------------------
def acoustic_op_problem(nt_data=151, dt_data=0.002, dt=0.001, sub=1, bc=(1, 2, 2, 2), seismotype=0, nx=50, nz=40):
    h = 0.01
    x = np.arange(nx)[:, None] * h
    z = np.arange(nz)[None, :] * h
    vp = 1.8 + 1.0 * z + 0 * x
    rho = 2.0 + 0.15 * np.sin(2 * np.pi * x / 0.45) * np.cos(2 * np.pi * z / 0.35)
    rx = np.stack([np.arange(5, nx - 5) * h, np.full(nx - 10, 0.05)], 1)
    prob = W.base_problem(nx, nz, h, h, nt_data, dt_data, [[0.23, 0.12], [0.31, 0.09]], [rx, rx[::2]], taper=0,
                          bc=bc, seismotype=seismotype, dt=dt, sub=sub)
    prob["acoustic"] = 1
    src = np.stack([W.ricker(nt_data, dt_data, 12.0), W.ricker(nt_data, dt_data, 9.0)])[None]
    return prob, np.stack([vp, rho]), src


@pytest.mark.parametrize("seismotype", [0, 1])
def test_acoustic_op_forward_is_the_propagation(orc64, seismotype):
    """At dt_data = dt the sinc resampler is the identity: the op's data are the
    propagation's receiver traces on the (K, 1/rho) model with the free-surface penalty of
    the default stiffness 1.05, times dt/dt_data = 1, and their time derivative Dt when
    seismotype = 1; vp above vpmax is clipped first."""
    prob, m, src = acoustic_op_problem(nt_data=301, dt_data=0.001, seismotype=seismotype)
    prob["mp"]["vpmax"] = 3.0
    m_hi = m.copy()
    m_hi[0, :5, :5] = 50.0  # clipped to vpmax = 3 by analyzeModel (courant 0.3)
    d, _ = orc64.apply_forward(orc64.make_op(prob), m_hi, src)
    m_clip = m_hi.copy()
    m_clip[0] = np.minimum(m_clip[0], prob["mp"]["vpmax"])
    model = np.stack([m_clip[1] * m_clip[0] ** 2, 1 / m_clip[1]])
    g = grid_struct(dict(prob, nt=301, dt=0.001, sub=1))
    nx, nz, h = prob["nx"], prob["nz"], prob["dx"]
    ref = []
    for s, rxz in enumerate(prob["rxz"]):
        inj = orc64.delta_m3(nx, nz, h, h, 0, 0, [prob["sxz"][s, 0]], [prob["sxz"][s, 1]])
        ext = orc64.delta_m3(nx, nz, h, h, 0, 0, rxz[:, 0], rxz[:, 1])
        ref.append(orc64.propagate_acoustic(g, model, inj, src[0, s:s + 1], ext=ext, alpha=A_FREE / 1.05))
    ref = np.concatenate(ref)
    assert np.abs(ref).max() < 1.0  # stable (the clip holds the corner at courant 0.3)
    if seismotype == 1:
        ref = orc64.Dt(False, ref, 0.001)
    assert d.shape == (1, ref.shape[0], 301)
    assert np.abs(d[0] - ref).max() <= 1e-9 * np.abs(ref).max()


def _acoustic_op_misfit(orc, prob, true, start, src):
    dt = prob["mp"]["dt_data"]
    op = orc.make_op(prob)
    dobs, _ = orc.apply_forward(op, true, src)
    d0, full = orc.apply_forward(op, start, src, keep_full=True)

    def J(m):
        r = orc.apply_forward(orc.make_op(prob), m, src)[0] - dobs
        return 0.5 * dt * (np.sum(r ** 2) - 0.5 * np.sum(r[..., 0] ** 2 + r[..., -1] ** 2))

    rw = (d0 - dobs) * dt
    rw[..., 0] *= 0.5
    rw[..., -1] *= 0.5
    return J, orc.apply_jacobianT(op, start, rw, full)


@pytest.mark.parametrize("seismotype,bc", [(0, (1, 3, 3, 3)), (1, (1, 3, 3, 3)), (0, (1, 2, 2, 2))])
def test_acoustic_op_gradient_finite_difference(orc64, seismotype, bc):
    """apply_jacobianT is the gradient of the trapezoid-rule misfit through the resampler
    (dt_data = 2 dt), the time derivative (seismotype 1), the H quadrature and the chain
    rule to (vp, rho): along vp (rho fixed) only the exact K gradient enters; along rho at
    fixed K (vp = sqrt(K/rho)) the chain rule's K and 1/rho terms both enter, with the
    1/rho term's narrow-versus-wide stencil difference (0.1-0.3 % here).  Absorbing sides
    take their impedance from the boundary cells' velocity, a dependence the reference's
    gradient leaves out (0.4 % along vp on this problem): with absorbing sides the
    perturbation is kept off the absorbing boundary cells, and the gradient is exact again."""
    from scipy.ndimage import gaussian_filter
    prob, true, src = acoustic_op_problem(seismotype=seismotype, bc=bc)
    prob["free_surface_stiffness"] = 2.0
    start = true.copy()
    start[0] = gaussian_filter(true[0], 4) * 0.97
    start[1] = gaussian_filter(true[1], 4) * 1.03
    J, g = _acoustic_op_misfit(orc64, prob, true, start, src)
    assert g.shape == (2,) + true.shape[1:]
    s = gaussian_filter(np.random.default_rng(3).standard_normal(true[0].shape), 3) * 0.01
    if bc[1] == 2:
        s[:1], s[-1:], s[:, -1:] = 0, 0, 0  # off the absorbing boundary cells
    K0 = start[1] * start[0] ** 2
    paths = {"vp": lambda e: np.stack([start[0] * (1 + e * s), start[1]]),
             "rho@K": lambda e: np.stack([np.sqrt(K0 / (start[1] * (1 + e * s))), start[1] * (1 + e * s)])}
    errs = {}
    for name, path in paths.items():
        fd = (J(path(0.1)) - J(path(-0.1))) / 0.2
        dm = (path(1e-6) - path(-1e-6)) / 2e-6
        errs[name] = (np.sum(g * dm) - fd) / fd
        print("seismotype %d bc %s %s: rel err %.2e" % (seismotype, bc, name, errs[name]))
    assert abs(errs["vp"]) < 1e-5
    assert abs(errs["rho@K"]) < 0.01
