"""The density term of the imaging condition collected per snapshot (by parts in time).

The engine images the density term of snapshot j after adjoint step it = nt-1-j as
<F_j, wn a_j-1 + wc a_j + wp a_j+1> (fwi2d_amd/csrc/engine.hip: sbp_image_args) instead
of the reference's sum over imaged snapshots idx = 0 .. nt-2 of
wrho_idx <a_idx, D2 F_idx> with D2 F_idx = F_idx+1 - 2 F_idx + F_idx-1 and the one-sided
F_1 - F_0 at idx = 0 (we_op.cpp:533-551).  This checks that the weight rule gives the
same sum for random fields, weights (a gdamp-like window) and every nt down to 2.
"""
import numpy as np
import pytest


def direct(F, a, w):
    nt = len(F)
    s = 0.0
    for idx in range(nt - 1):  # imaged snapshots 0 .. nt-2 (sub = 1)
        d2 = F[1] - F[0] if idx == 0 else F[idx + 1] - 2 * F[idx] + F[idx - 1]
        s += w[idx] * np.vdot(a[idx], d2)
    return s


def by_parts(F, a, w):
    """The weight rule of sbp_image_args: snapshot j meets a_j-1, a_j, a_j+1."""
    nt = len(F)
    last = nt - 2
    s = 0.0
    for j in range(nt):
        wn = w[j - 1] if 1 <= j and j - 1 <= last else 0.0
        wc = -2 * w[j] if 1 <= j <= last else (-w[0] if j == 0 else 0.0)
        wp = w[j + 1] if j + 1 <= last else 0.0
        acc = np.zeros_like(F[0])
        if wn:
            acc = acc + wn * a[j - 1]
        if wc:
            acc = acc + wc * a[j]
        if wp:
            acc = acc + wp * a[j + 1]
        s += np.vdot(F[j], acc)
    return s


@pytest.mark.parametrize("nt", [2, 3, 4, 7, 40])
def test_density_term_by_parts(nt):
    rng = np.random.default_rng(nt)
    F = rng.standard_normal((nt, 5, 3))
    a = rng.standard_normal((nt, 5, 3))
    w = np.cos(np.linspace(0, 1.2, nt)) ** 2  # a varying window of weights
    d, b = direct(F, a, w), by_parts(F, a, w)
    assert abs(d - b) <= 1e-12 * (abs(d) + 1.0), (nt, d, b)


def test_by_parts_catches_a_dropped_endpoint():
    """A rule that forgets the one-sided first snapshot (-w0 a0 on F_0) must differ."""
    nt = 6
    rng = np.random.default_rng(1)
    F, a = rng.standard_normal((nt, 4)), rng.standard_normal((nt, 4))
    w = np.ones(nt)
    wrong = by_parts(F, a, w) + w[0] * np.vdot(F[0], a[0])  # the -w0 a0 term of F_0 dropped
    assert abs(wrong - direct(F, a, w)) > 1e-6
