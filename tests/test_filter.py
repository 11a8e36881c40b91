"""The objective's 1-D filter (nlls_fwi_eco's filter argument, src/include/optimization.hpp:
526-539, 688-696, 846-854), on the CPU.

Oracle pins (oracle/oracle.py: conv1dnd, zero_phase, minimum_phase):
  * conv1dnd (src/cpp/operator.cpp:2092-2137) is a slice of the full discrete convolution
    (numpy.convolve) at the shift floor((nf+1)/2) - 1 (centred) or 0, for odd, even and
    over-long filters; its adjoint passes the dot-product test.
  * zero_phase (operator.cpp:2197-2250) keeps the amplitude spectrum and returns a
    symmetric filter; a symmetric filter with a positive spectrum (a sampled Gaussian) comes
    back unchanged, and a delayed copy of it comes back centred.
  * minimum_phase (operator.cpp:2252-2320) keeps the amplitude spectrum (to the 1e-7 of
    its log guard); a minimum-phase filter (1 - a z^-1, |a| < 1) comes back unchanged, and
    its maximum-phase mirror (-a + z^-1) comes back as the minimum-phase filter; the result
    has its energy front-loaded at least as much as other filters of the same spectrum.
Product (host C++ through the C ABI, no device): fwi2d_filter_prepare and
fwi2d_filter_apply equal the oracle's, and nlls_fwi_eco's data norm includes the filter.
"""
import os

import numpy as np
import pytest

from fwi2d_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "fwi2d_amd", "libfwi2d_hip.so")
needs_lib = pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")


@pytest.mark.parametrize("nf,centered", [(7, True), (8, True), (8, False), (1, True), (57, True), (57, False)])
def test_conv1dnd_is_a_slice_of_the_full_convolution(orc64, nf, centered):
    rng = np.random.default_rng(nf)
    nt = 40
    x = rng.standard_normal((3, nt))
    f = rng.standard_normal(nf)
    y = orc64.conv1dnd(x, f, centered)
    sh = (nf + 1) // 2 - 1 if centered else 0
    for k in range(3):
        full = np.convolve(x[k], f)
        assert np.allclose(y[k], full[sh:sh + nt], rtol=0, atol=1e-13)
    v = rng.standard_normal((3, nt))
    assert np.sum(y * v) == pytest.approx(np.sum(x * orc64.conv1dnd(v, f, centered, adjoint=True)), rel=1e-12)


def gaussian(n, c, s):
    k = np.arange(n)
    return np.exp(-0.5 * ((k - c) / s) ** 2)


def test_zero_phase(orc64):
    rng = np.random.default_rng(1)
    for n in (31, 30):
        f = rng.standard_normal(n)
        z = orc64.zero_phase(f)
        assert z.size == n + (n % 2 == 0)
        fp = np.concatenate([f, np.zeros(z.size - n)])
        assert np.allclose(np.abs(np.fft.rfft(z)), np.abs(np.fft.rfft(fp)), rtol=1e-10, atol=1e-12)
        assert np.allclose(z, z[::-1], atol=1e-12)
    g = gaussian(41, 20, 2.0)
    assert np.allclose(orc64.zero_phase(g), g, atol=1e-12)
    # a delayed Gaussian (linear phase) comes back centred
    assert np.allclose(orc64.zero_phase(gaussian(41, 13, 2.0)), g, atol=1e-9)


def test_minimum_phase(orc64):
    n, a = 101, 0.5
    mn = np.zeros(n)
    mn[0], mn[1] = 1.0, -a
    mx = np.zeros(n)
    mx[0], mx[1] = -a, 1.0
    assert np.allclose(orc64.minimum_phase(mn), mn, atol=1e-6)
    assert np.allclose(orc64.minimum_phase(mx), mn, atol=1e-6)
    rng = np.random.default_rng(2)
    f = rng.standard_normal(64)
    m = orc64.minimum_phase(f)
    fp = np.concatenate([f, [0.0]])
    assert np.allclose(np.abs(np.fft.rfft(m)), np.abs(np.fft.rfft(fp)), rtol=1e-5, atol=1e-6)
    # energy front-loading: partial energies of the minimum-phase filter dominate (to the
    # cepstral aliasing of the finite transform length, 1.2e-5 of the energy here)
    for other in (fp, fp[::-1], orc64.zero_phase(f)):
        assert np.all(np.cumsum(m ** 2) >= np.cumsum(other ** 2) - 1e-4 * np.sum(m ** 2))
    assert np.sum(m[:8] ** 2) > 0.7 * np.sum(m ** 2) > np.sum(fp[:8] ** 2)


@needs_lib
@pytest.mark.parametrize("phase", ["default", "zero", "minimum"])
@pytest.mark.parametrize("nf", [1, 24, 25, 60, 301])
def test_host_filter_matches_oracle(orc64, phase, nf):
    from fwi2d_amd import lib
    rng = np.random.default_rng(nf)
    f = rng.standard_normal(nf).astype(np.float32)
    fh, ch = lib.filter_prepare(f, phase)
    fo, co = orc64.prepare_filter(f.astype(np.float64), phase)
    assert ch == co and fh.size == fo.size
    assert np.allclose(fh, fo, rtol=0, atol=1e-10 * np.abs(fo).max())
    x = rng.standard_normal((5, 60))
    for adjoint in (False, True):
        assert np.allclose(lib.filter_apply(fh, x, ch, adjoint), orc64.conv1dnd(x, fo, co, adjoint), rtol=0,
                           atol=1e-12 * np.abs(x).max() * np.abs(fo).sum())


@needs_lib
def test_filter_arguments_are_checked():
    from fwi2d_amd import lib
    n = lib.C.c_int()
    with pytest.raises(lib.Fwi2dError, match="filter phase"):
        lib._check(lib.load().fwi2d_filter_prepare(np.ones(3, np.float32).ctypes.data, 3, 7,
                                                   np.zeros(4).ctypes.data, lib.C.byref(n)))
    with pytest.raises(lib.Fwi2dError, match="at least one sample"):
        lib._check(lib.load().fwi2d_filter_prepare(np.ones(3, np.float32).ctypes.data, 0, 0,
                                                   np.zeros(4).ctypes.data, lib.C.byref(lib.C.c_int())))


@needs_lib
@pytest.mark.parametrize("phase", ["default", "zero", "minimum"])
def test_data_norm_includes_the_filter(orc64, phase):
    """nlls_fwi_eco's dnorm = dt |C W d|^2: the constructor filters the observed data after
    the weights (optimization.hpp:524-539, 569-574).  Creating the objective touches no
    device."""
    from fwi2d_amd import lib
    nx, nz, h = 40, 30, 0.01
    rx = np.stack([np.arange(2, nx - 2, 4) * h, np.full(len(range(2, nx - 2, 4)), 0.02)], 1)
    prob = W.base_problem(nx, nz, h, h, 101, 0.002, [[0.1, 0.02]], [rx], taper=5, dt=0.001, sub=-1)
    op = lib.WeOp(prob, W.vertical_force(W.ricker(101, 0.002, 12.0), 1))
    rng = np.random.default_rng(3)
    d = rng.standard_normal((2, op.nr, op.nt_data)).astype(np.float32)
    w = rng.uniform(0.5, 2.0, d.shape).astype(np.float32)
    f = rng.standard_normal(9).astype(np.float32)
    fo, co = orc64.prepare_filter(f.astype(np.float64), phase)
    cd = orc64.conv1dnd(d.astype(np.float64) * w.astype(np.float64), fo, co)
    obj = lib.Nlls(op, d, w=w, normalize=True, filter=f, filter_phase=phase)
    assert obj.dnorm() == pytest.approx(0.002 * np.sum(cd ** 2), rel=1e-12)
    obj.close()
    op.close()
