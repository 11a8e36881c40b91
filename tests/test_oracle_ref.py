"""Oracle vs the reference's own SBP code compiled from its sources (oracle/_ref).

The reference's operator templates (spatial_operators.hpp) and the non-ISPC part of
spatial_operators.cpp build as they lie; each oracle routine below is compared with
the compiled reference routine on seeded random fields (FP64 on both sides).
"""
import numpy as np
import pytest

SIDES = {0: "top", 1: "bottom", 2: "left", 3: "right"}


@pytest.fixture(scope="module")
def fields():
    rng = np.random.default_rng(11)
    nx, nz = 23, 19
    f = {k: rng.standard_normal((nx, nz)) for k in ("u", "in0", "in1", "in2", "out")}
    f.update({k: 1.0 + rng.random((nx, nz)) for k in ("c", "f1", "f2", "f3")})
    lam, mu, rho = 1 + rng.random((nx, nz)), 1 + rng.random((nx, nz)), 1 + rng.random((nx, nz))
    f["model"] = np.stack([lam, mu, rho, rng.random((nx, nz)), 0.2 * rng.random((nx, nz))])
    return f


@pytest.mark.parametrize("name", ["Dzz_var", "Dxx_var"])
def test_d2var_matches_reference(orc64, refsbp, fields, name):
    a = getattr(orc64, name)(fields["u"], fields["c"], 0.013)
    b = getattr(refsbp, name)(fields["u"], fields["c"], 0.013)
    assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max()


def test_d2var_unit_coefficient_is_reference_constant_d2(orc64, refsbp):
    n = 20
    e = np.eye(n)
    a = orc64.Dzz_var(e, np.ones((n, n)), 1.0)
    b = refsbp.Dzz(e, 1.0)
    assert np.abs(a - b).max() < 1e-13


@pytest.mark.parametrize("side", [0, 1, 2, 3])
@pytest.mark.parametrize("absorbing", [False, True])
def test_sat_matches_reference(orc64, refsbp, fields, side, absorbing):
    f = fields
    in2 = f["in2"] if absorbing else None
    f3 = f["f3"] if absorbing else None
    a = orc64.esat(side, f["in0"], f["in1"], in2, f["out"], f["f1"], f["f2"], f3, 0.011, 0.013, 0.001)
    b = refsbp.esat(side, f["in0"], f["in1"], in2, f["out"], f["f1"], f["f2"], f3, 0.011, 0.013, 0.001)
    assert np.abs(b - f["out"]).max() > 0
    assert np.abs(a - b).max() <= 1e-13 * np.abs(b - f["out"]).max()


@pytest.mark.parametrize("side", [0, 1, 2, 3])
def test_sat_normal_complement_matches_reference(orc64, refsbp, fields, side):
    f = fields
    a = orc64.esat_dn(side, f["in0"], f["out"], f["f1"], 0.011, 0.013)
    b = refsbp.esat_dn(side, f["in0"], f["out"], f["f1"], 0.011, 0.013)
    assert np.abs(a - b).max() <= 1e-13 * np.abs(b - f["out"]).max()


def test_tangential_d1_through_reference_sat(orc64, refsbp, fields):
    """With f2 = 0 and f1 = 1 the reference's top SAT returns the x-derivative of row 0:
    out = (1/(dz h0)) (Dx in0)_row0 -- this pins the oracle's D1 (closure and interior)
    against reference code."""
    f = fields
    nx, nz = f["u"].shape
    dx, dz = 0.011, 0.013
    ones, zeros = np.ones((nx, nz)), np.zeros((nx, nz))
    b = refsbp.esat(0, f["u"], zeros, None, zeros, ones, zeros, None, dx, dz, 0.001)
    d1 = orc64.Dx(f["u"], dx)
    h0 = 17.0 / 48
    np.testing.assert_allclose(b[:, 0] * dz * h0, d1[:, 0], rtol=0, atol=1e-12 * np.abs(d1[:, 0]).max())


@pytest.mark.parametrize("vti", [0, 1])
@pytest.mark.parametrize("flags", [(1, 1, 1, 1), (0, 1, 1, 1), (1, 0, 1, 0), (0, 0, 0, 1), (1, 1, 0, 0), (0, 0, 1, 0)])
def test_boundary_scaling_matches_reference(orc64, refsbp, fields, vti, flags):
    f = fields
    m = f["model"][: 5 if vti else 3]
    a = orc64.scale_boundaries(f["in0"], f["in1"], m, vti, 0.011, 0.013, 0.001, *flags)
    b = refsbp.scale_boundaries(f["in0"], f["in1"], m, vti, 0.011, 0.013, 0.001, *flags)
    for i in range(2):
        assert np.abs(a[i] - b[i]).max() <= 1e-15 * np.abs(b[i]).max()


@pytest.mark.parametrize("axis_x", [0, 1])
@pytest.mark.parametrize("inv", [0, 1])
def test_quadrature_matches_reference(orc64, refsbp, fields, axis_x, inv):
    a = orc64.applyH(fields["u"], 0.012, axis_x, inv)
    b = refsbp.applyH(fields["u"], 0.012, axis_x, inv)
    assert np.abs(a - b).max() <= 1e-14 * np.abs(b).max()
