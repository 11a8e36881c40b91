"""Host logic of the least-squares objective (fwi2d_nlls_*), without a GPU: the
constructor's checks and data norm (nlls_fwi / nlls_fwi_eco constructors,
src/include/optimization.hpp:309-344, 482-590).  Creating the operator and the objective
touches no device; evaluation needs the GPU and is covered by the GPU parity tests."""
import os

import numpy as np
import pytest

from fwi2d_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "fwi2d_amd", "libfwi2d_hip.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")


@pytest.fixture(scope="module")
def op():
    from fwi2d_amd import lib
    nx, nz, h = 40, 30, 0.01
    rx = np.stack([np.arange(2, nx - 2, 4) * h, np.full(len(range(2, nx - 2, 4)), 0.02)], 1)
    prob = W.base_problem(nx, nz, h, h, 101, 0.002, [[0.1, 0.02], [0.3, 0.05]], [rx, rx[::2]], taper=5,
                          dt=0.001, sub=-1)
    src = W.vertical_force(W.ricker(101, 0.002, 12.0), 2)
    o = lib.WeOp(prob, src)
    yield o
    o.close()


def test_data_norm(op):
    from fwi2d_amd import lib
    rng = np.random.default_rng(0)
    d = rng.standard_normal((2, op.nr, op.nt_data)).astype(np.float32)
    w = rng.uniform(0.5, 2.0, d.shape).astype(np.float32)
    wd = d.astype(np.float64) * w.astype(np.float32)
    assert lib.Nlls(op, d).dnorm() == 1.0
    assert lib.Nlls(op, d, normalize=True).dnorm() == pytest.approx(0.002 * np.sum(d.astype(np.float64) ** 2),
                                                                    rel=1e-12)
    # the weights multiply the data before the norm
    assert lib.Nlls(op, d, w=w, normalize=True).dnorm() == pytest.approx(0.002 * np.sum(wd ** 2), rel=1e-12)
    # trace normalisation: every trace has unit norm, dnorm = dt * ntraces
    assert lib.Nlls(op, d, normalize=True, trace_normalize=True).dnorm() == pytest.approx(0.002 * 2 * op.nr,
                                                                                          rel=1e-12)
    # integration: the trapezoid running sum of the data
    tw = np.full(op.nt_data, 0.002)
    tw[0] = tw[-1] = 0.001
    di = np.cumsum(d.astype(np.float64) * tw, axis=-1)
    di[..., -1] = di[..., -2] + 0.001 * d[..., -1]
    assert lib.Nlls(op, d, normalize=True, integrate=True).dnorm() == pytest.approx(0.002 * np.sum(di ** 2),
                                                                                    rel=1e-12)
    # all-zero data: dnorm falls back to 1
    assert lib.Nlls(op, np.zeros_like(d), normalize=True).dnorm() == 1.0
    assert lib.Nlls(op, d).residual() is None


def test_argument_checks(op):
    from fwi2d_amd import lib
    d = np.zeros((2, op.nr, op.nt_data), np.float32)
    w = np.ones_like(d)
    w[1, 3, 7] = -1e-3
    with pytest.raises(lib.Fwi2dError, match="Data weights  must be non-negative"):
        lib.Nlls(op, d, w=w)
    mask = np.ones((3, 40, 30), np.float32)
    mask[2, 5, 5] = -1.0
    with pytest.raises(lib.Fwi2dError, match="Gradient mask must be non-negative"):
        lib.Nlls(op, d, gmask=mask)
    mask[2, 5, 5] = np.nan
    with pytest.raises(lib.Fwi2dError, match="Gradient mask must be non-negative"):
        lib.Nlls(op, d, gmask=mask)
    L = lib.load()
    assert L.fwi2d_nlls_create(None, None, None, None, 0, None) != 0
    assert b"null argument" in L.fwi2d_last_error()
    h = lib.C.c_void_p()
    assert L.fwi2d_nlls_create(op._h, d.ctypes.data_as(lib.C.c_void_p), None, None, 8, lib.C.byref(h)) != 0
    assert b"unknown objective flags" in L.fwi2d_last_error()


def test_das_and_moment_tensor_arguments_are_checked():
    """DAS channels must keep both gauge ends inside the model (we_op.cpp:134-148), need
    their dips, and moment-tensor sources come as three traces per shot."""
    from fwi2d_amd import lib
    nx, nz, h = 40, 30, 0.01
    rx = np.stack([np.arange(2, nx - 2, 4) * h, np.full(len(range(2, nx - 2, 4)), 0.05)], 1)
    prob = W.base_problem(nx, nz, h, h, 101, 0.002, [[0.1, 0.1]], [rx], taper=5, dt=0.001, sub=-1)
    w = W.ricker(101, 0.002, 12.0)
    prob_das = dict(prob, gl=0.05, rdip=[np.zeros(len(rx))])
    with pytest.raises(lib.Fwi2dError, match="half of the gauge length"):
        lib.WeOp(prob_das, W.vertical_force(w, 1))  # the first channel's gauge starts at x = -0.005 km
    ok = dict(prob_das, rxz=[rx[1:]], rdip=[np.zeros(len(rx) - 1)])
    lib.WeOp(ok, W.vertical_force(w, 1)).close()
    with pytest.raises(AssertionError):
        lib.WeOp(dict(prob, mt=1), W.vertical_force(w, 1))  # [2][ns][nt] given, [3][ns][nt] expected
    lib.WeOp(dict(prob, mt=1), np.stack([w[None], w[None], 0 * w[None]])).close()


def test_acoustic_operator_arguments():
    """The acoustic operator (nparam 2, nl_we_op_a): one source trace per shot
    (analyzeWavelet's acoustic copy, we_op.cpp:178-194) and one data component; the
    objective's data and gradient mask take those shapes; moment tensors and DAS are
    switched off as we_op.cpp:13-20 does; parameter counts other than 2, 3 and 5 fail."""
    from fwi2d_amd import lib
    nx, nz, h = 40, 30, 0.01
    rx = np.stack([np.arange(2, nx - 2, 4) * h, np.full(len(range(2, nx - 2, 4)), 0.05)], 1)
    prob = W.base_problem(nx, nz, h, h, 101, 0.002, [[0.1, 0.1], [0.2, 0.1]], [rx, rx], taper=5, dt=0.001, sub=-1)
    prob["acoustic"] = 1
    w = W.ricker(101, 0.002, 12.0)
    src = np.stack([w, w])[None].astype(np.float32)
    with pytest.raises(AssertionError):
        lib.WeOp(prob, W.vertical_force(w, 2))  # [2][ns][nt]: a force, not an acoustic source
    op = lib.WeOp(dict(prob, mt=1, gl=0.05), src)  # mt and gl ignored, as the reference does
    assert op.np == 2 and op.ncd == 1
    d = np.random.default_rng(1).standard_normal((1, op.nr, op.nt_data)).astype(np.float32)
    assert lib.Nlls(op, d, normalize=True).dnorm() == pytest.approx(0.002 * np.sum(d.astype(np.float64) ** 2),
                                                                    rel=1e-12)
    with pytest.raises(AssertionError):
        lib.Nlls(op, np.zeros((2, op.nr, op.nt_data), np.float32))
    mask = np.ones((2, nx, nz), np.float32)
    mask[1, 3, 3] = -1.0
    with pytest.raises(lib.Fwi2dError, match="Gradient mask must be non-negative"):
        lib.Nlls(op, d, gmask=mask)
    op.close()
    d4 = lib.op_desc_from(prob)
    d4.nparam = 4
    sxz = np.ascontiguousarray(prob["sxz"])
    nrec = np.array([len(rx), len(rx)], np.int32)
    rxz = np.ascontiguousarray(np.concatenate([rx, rx]))
    d4.ns = 2
    d4.sxz = sxz.ctypes.data_as(lib.C.POINTER(lib.C.c_double))
    d4.nrec = nrec.ctypes.data_as(lib.C.POINTER(lib.C.c_int))
    d4.rxz = rxz.ctypes.data_as(lib.C.POINTER(lib.C.c_double))
    h4 = lib.C.c_void_p()
    L = lib.load()
    assert L.fwi2d_op_create(lib.C.byref(d4), src.ctypes.data_as(lib.C.c_void_p), lib.C.byref(h4)) != 0
    assert b"two (Vp, rho), three (Vp, Vs, rho) or five" in L.fwi2d_last_error()
