"""Pins of the oracle's moment-tensor and DAS injectors (ddelta_m3, dipole_m3,
dipole_to_strain: src/cpp/injector.cpp:114-378, src/cpp/misc.cpp:18-43), on the CPU.

  * ddelta_m3's derivative weights are minus the derivative of delta_m3's weights with
    respect to the point's position (exactly: the weights are quadratic in the position,
    so a central difference is exact), for injection (H-scaled) and extraction; with the
    H weights they reproduce -f'(x_s) for f = 1, x, x^2.
  * The moment-tensor source is the equivalent body force f_i = -d_j (M_ij delta(x - xs))
    = M_ij d/dxs_j delta(x - xs): its seismograms equal central differences, in the source
    position, of point-force seismograms (Mxx <-> d/dxs of the x force, Mzz <-> d/dzs of
    the z force, Mxz <-> d/dzs of the x force + d/dxs of the z force), to rounding.  A
    swapped component or sign is off by O(1).
  * An explosive source (Mxx = Mzz, Mxz = 0) in a homogeneous full space radiates no S
    wave: its displacement at a receiver is along the source-receiver direction (exactly
    on the grid axes and diagonals, to the grid anisotropy elsewhere), where a pure shear
    source (Mxx = -Mzz) is not.
  * dipole_m3's footprints are delta_m3's at the two ends of the gauge (x -/+ gl/2 cos dip,
    z -/+ gl/2 sin dip); a DAS channel records the difference of point receivers at the two
    ends projected on the fiber over gl (dipole_to_strain); dipole_to_strain's adjoint
    passes the dot-product test.
"""
import numpy as np
import pytest

from fwi2d_amd import workloads as W


def h_weight(n, i):
    c = [17 / 48, 59 / 48, 43 / 48, 49 / 48]
    return c[i] if i < 4 else (c[n - 1 - i] if i >= n - 4 else 1.0)


@pytest.mark.parametrize("x,z", [(0.237, 0.143), (0.012, 0.41), (0.3449, 0.0031)])
def test_ddelta_weights_are_position_derivatives(orc64, x, z):
    nx, nz, h = 40, 45, 0.01
    xi, zi, xw, zw = orc64.ddelta_m3(nx, nz, h, h, 0.0, 0.0, [x], [z])
    eps = 1e-4 * h
    for axis in (0, 1):
        lp = [x + eps * (axis == 0)], [z + eps * (axis == 1)]
        lm = [x - eps * (axis == 0)], [z - eps * (axis == 1)]
        dp, dm = orc64.delta_m3(nx, nz, h, h, 0.0, 0.0, *lp), orc64.delta_m3(nx, nz, h, h, 0.0, 0.0, *lm)
        w = (xw, zw)[axis][0]
        ind = 2 + axis
        fd = (dp[ind][0] - dm[ind][0]) / (2 * eps)
        assert np.allclose(w[:, 3:], -fd, rtol=1e-6, atol=1e-6 * np.abs(w[:, 3:]).max())
        assert np.array_equal((xw, zw)[axis][0][:, :3], dp[ind][0] * 0 + (orc64.delta_m3(
            nx, nz, h, h, 0.0, 0.0, [x], [z])[ind][0]))
    # moments with the H weights: sum h H_a d_a f(x_a) = -f'(x_s) for f = 1, x, x^2
    xa = xi[0] * h
    Ha = np.array([h_weight(nx, i) for i in xi[0]]) * h
    for f, fp in ((lambda s: 1 + 0 * s, 0.0), (lambda s: s, 1.0), (lambda s: s * s, 2 * x)):
        assert np.sum(Ha * xw[0, 0, 3:] * f(xa)) == pytest.approx(-fp, abs=1e-10)
        assert np.sum(xw[0, 1, 3:] * f(xa)) == pytest.approx(-fp, abs=1e-10)


def homogeneous(nx=90, nz=90, h=0.01, vp=2.0, vs=1.15, rho=2.0):
    lam, mu = rho * (vp ** 2 - 2 * vs ** 2), rho * vs ** 2
    model = np.stack([np.full((nx, nz), lam), np.full((nx, nz), mu), np.full((nx, nz), rho)])
    prob = dict(nx=nx, nz=nz, dx=h, dz=h, bc_top=2, bc_bottom=2, bc_left=2, bc_right=2, taper_top=15,
                taper_bottom=15, taper_left=15, taper_right=15, taper_strength=0.05, vti=0)
    return prob, model


def op_problem(mt, rxz, nt_data=181, dt=0.001, sx=(0.4531, 0.4467), gl=0.0, dips=None):
    prob, model = homogeneous()
    p = W.base_problem(prob["nx"], prob["nz"], prob["dx"], prob["dz"], nt_data, dt, [list(sx)], [rxz], taper=15,
                       bc=(2, 2, 2, 2), seismotype=0, dt=dt, sub=0)
    p["mt"] = mt
    if gl > 0:
        p["gl"] = gl
        p["rdip"] = [np.asarray(dips, np.float64)]
    vp = np.sqrt((model[0] + 2 * model[1]) / model[2])
    vs = np.sqrt(model[1] / model[2])
    return p, np.stack([vp, vs, model[2]])


def test_moment_tensor_is_the_equivalent_body_force(orc64):
    rxz = np.array([[0.20, 0.25], [0.65, 0.41], [0.47, 0.70], [0.30, 0.62]])
    nt_data, dt = 181, 0.001
    w = W.ricker(nt_data, dt, 25.0)
    sx = (0.4531, 0.4467)
    p_mt, m = op_problem(1, rxz, nt_data, dt, sx)
    eps = 1e-3 * 0.01

    def force(comp, dx=0.0, dz=0.0):
        p, _ = op_problem(0, rxz, nt_data, dt, (sx[0] + dx, sx[1] + dz))
        src = np.zeros((2, 1, nt_data))
        src[comp, 0] = w
        return orc64.apply_forward(orc64.make_op(p), m, src)[0]

    def d_ds(comp, axis):
        return (force(comp, *(eps * (axis == 0), eps * (axis == 1))) -
                force(comp, *(-eps * (axis == 0), -eps * (axis == 1)))) / (2 * eps)
    cases = {"mxx": d_ds(0, 0), "mzz": d_ds(1, 1), "mxz": d_ds(0, 1) + d_ds(1, 0)}
    for k, (name, ref) in enumerate(cases.items()):
        src = np.zeros((3, 1, nt_data))
        src[k, 0] = w  # [mxx, mzz, mxz]
        out = orc64.apply_forward(orc64.make_op(p_mt), m, src)[0]
        err = np.linalg.norm(out - ref) / np.linalg.norm(ref)
        print("%s: moment tensor vs d/ds of point forces rel err %.2e" % (name, err))
        assert err < 1e-6


def test_explosive_source_radiates_no_shear(orc64):
    rxz = np.array([[0.2, 0.45], [0.45, 0.2], [0.27, 0.27], [0.65, 0.62]])
    nt_data, dt = 260, 0.001
    w = W.ricker(nt_data, dt, 20.0)
    sx = (0.45, 0.45)
    p, m = op_problem(1, rxz, nt_data, dt, sx)
    for mzz, expect_radial in ((1.0, True), (-1.0, False)):
        src = np.zeros((3, 1, nt_data))
        src[0, 0], src[1, 0] = w, mzz * w
        out = orc64.apply_forward(orc64.make_op(p), m, src)[0]
        ratios = []
        for r in range(len(rxz)):
            d = rxz[r] - np.array(sx)
            d /= np.linalg.norm(d)
            radial = d[0] * out[0, r] + d[1] * out[1, r]
            tang = -d[1] * out[0, r] + d[0] * out[1, r]
            ratios.append(np.linalg.norm(tang) / max(np.linalg.norm(radial), 1e-300))
        print("mxx=1 mzz=%g: tangential / radial amplitude per receiver %s" % (mzz, ["%.2e" % q for q in ratios]))
        if expect_radial:  # exact by symmetry on the axes and diagonals; 3.6e-2 off them (grid
            # anisotropy at 10 cells per P wavelength)
            assert max(ratios[:3]) < 1e-12 and max(ratios) < 6e-2
        else:  # a pure shear source (mxx = -mzz) has S waves: not radial off its nodal lines
            assert max(ratios) > 0.3


def test_dipole_footprints_and_strain(orc64):
    nx, nz, h, gl = 40, 45, 0.01, 0.02
    x = np.array([0.171, 0.2, 0.25])
    z = np.array([0.113, 0.2, 0.31])
    dip = np.array([0.3, 0.0, 1.2])
    xi, zi, xw, zw = orc64.dipole_m3(nx, nz, h, h, 0.0, 0.0, x, z, dip, gl)
    for e, sgn in ((0, -1), (1, 1)):
        ex, ez = x + sgn * gl / 2 * np.cos(dip), z + sgn * gl / 2 * np.sin(dip)
        dxi, dzi, dxw, dzw = orc64.delta_m3(nx, nz, h, h, 0.0, 0.0, ex, ez)
        assert np.array_equal(xi[:, 3 * e:3 * e + 3], dxi) and np.array_equal(zi[:, 3 * e:3 * e + 3], dzi)
        assert np.allclose(xw[:, 0, 3 * e:3 * e + 3], dxw[:, 0], rtol=1e-14)
        assert np.allclose(xw[:, 1, 3 * e:3 * e + 3], dxw[:, 1], rtol=1e-14)
        assert np.allclose(zw[:, 1, 3 * e:3 * e + 3], dzw[:, 1], rtol=1e-14)
    rng = np.random.default_rng(0)
    t = rng.standard_normal((2, 3, 50))
    s = orc64.dipole_to_strain(False, t, dip, gl)
    assert np.allclose(s[0], (np.cos(dip)[:, None] * t[0] + np.sin(dip)[:, None] * t[1]) / gl) and np.all(s[1] == 0)
    v = rng.standard_normal((2, 3, 50))
    v[1] = 0
    assert np.sum(s * v) == pytest.approx(np.sum(t * orc64.dipole_to_strain(True, v, dip, gl)), rel=1e-12)


def test_das_channel_is_a_projected_difference_of_point_receivers(orc64):
    gl = 0.03
    rxz = np.array([[0.25, 0.3], [0.6, 0.55], [0.4, 0.7]])
    dips = np.array([0.2, 1.1, -0.7])
    nt_data, dt = 181, 0.001
    w = W.ricker(nt_data, dt, 20.0)
    src = np.stack([0.3 * w, w])[:, None, :]
    p_das, m = op_problem(0, rxz, nt_data, dt, gl=gl, dips=dips)
    das = orc64.apply_forward(orc64.make_op(p_das), m, src)[0]
    ends = np.concatenate([rxz - gl / 2 * np.stack([np.cos(dips), np.sin(dips)], 1),
                           rxz + gl / 2 * np.stack([np.cos(dips), np.sin(dips)], 1)])
    p_pt, _ = op_problem(0, ends, nt_data, dt)
    pt = orc64.apply_forward(orc64.make_op(p_pt), m, src)[0]
    n = len(rxz)
    diff = pt[:, n:] - pt[:, :n]
    ref = (np.cos(dips)[:, None] * diff[0] + np.sin(dips)[:, None] * diff[1]) / gl
    assert np.allclose(das[0], ref, rtol=0, atol=1e-12 * np.abs(ref).max()) and np.all(das[1] == 0)
