"""Moment-tensor sources (par.mt, ddelta_m3) and DAS receivers (par.gl, dipole_m3 +
dipole_to_strain) through the operator level of the HIP path, against the FP64 oracle:
seismograms within 1e-4 and (vp, vs, rho) gradients within 1e-4, isotropic and VTI, with
resampling and particle-velocity (strain-rate) data; the batch of shots equals each shot
alone bit for bit."""
import numpy as np
import pytest

from fwi2d_amd import workloads as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def lib():
    import fwi2d_amd
    return fwi2d_amd


def case(mt, gl, vti=False, seismotype=1):
    nx, nz, h = 110, 70, 0.01
    nt_data, dt_data = 251, 0.002
    true = W.add_ellipses(W.gradient_model(nx, nz, h, h, seed=3), h, h, seed=4)
    if vti:
        true = np.concatenate([true, W.vti_fields(nx, nz, h, h)], 0)
    start = W.smooth_model(true, h, h, 0.1)
    rng = np.random.default_rng(5)
    rx = np.stack([np.arange(4, nx - 4, 3) * h, np.full(len(range(4, nx - 4, 3)), 0.05)], 1)
    rx2 = np.stack([np.full(12, 0.7), np.linspace(0.1, 0.6, 12)], 1)  # a vertical well
    prob = W.base_problem(nx, nz, h, h, nt_data, dt_data, [[0.3, 0.22], [0.8, 0.35]], [rx, rx2], vti=int(vti),
                          taper=15, seismotype=seismotype, dt=0.001, sub=-1)
    w = W.ricker(nt_data, dt_data, 12.0)
    if mt:
        src = np.stack([np.tile(w * m, (2, 1)) for m in (1.0, 0.6, -0.4)])  # mxx, mzz, mxz
        prob["mt"] = 1
    else:
        src = W.vertical_force(w, 2)
    if gl > 0:
        prob["gl"] = gl
        prob["rdip"] = [rng.uniform(-0.3, 0.3, len(rx)), np.full(len(rx2), np.pi / 2)]
    return prob, true, start, src.astype(np.float32)


@pytest.mark.parametrize("mt,gl,vti", [(1, 0.0, False), (0, 0.03, False), (1, 0.03, False), (1, 0.03, True)])
def test_operator_level_injectors_match_oracle(lib, orc64, mt, gl, vti):
    prob, true, start, src = case(mt, gl, vti)
    dobs, _ = orc64.apply_forward(orc64.make_op(prob), true, src)
    d0, full = orc64.apply_forward(orc64.make_op(prob), start, src, keep_full=True)
    op = lib.WeOp(prob, src)
    d0k = op.apply_forward(start)
    err_d = rel(d0k, d0)
    print("mt=%d gl=%g vti=%s: forward rel err %.3e" % (mt, gl, vti, err_d))
    assert err_d <= 1e-4
    if gl > 0:
        assert np.all(d0k[1] == 0)  # strain (rate) in component 0 only
    r = d0 - dobs
    rw = r * prob["mp"]["dt_data"]
    rw[..., 0] *= 0.5
    rw[..., -1] *= 0.5
    gref = orc64.apply_jacobianT(orc64.make_op(prob), start, rw, full)
    gk = op.apply_jacobianT(start, rw)
    for c in range(3):
        err = rel(gk[c], gref[c])
        print("mt=%d gl=%g vti=%s: gradient component %d rel err %.3e" % (mt, gl, vti, c, err))
        assert err <= 1e-4
    op.close()


def test_batch_of_moment_tensors_and_das_equals_single_shots(lib):
    prob, true, start, src = case(1, 0.03)
    op = lib.WeOp(prob, src)
    d = op.apply_forward(start)
    r0 = 0
    for s in range(2):
        ps = dict(prob, sxz=prob["sxz"][s:s + 1], rxz=prob["rxz"][s:s + 1], rdip=prob["rdip"][s:s + 1])
        one = lib.WeOp(ps, src[:, s:s + 1])
        n = len(prob["rxz"][s])
        assert np.array_equal(one.apply_forward(start), d[:, r0:r0 + n])
        r0 += n
        one.close()
    op.close()
