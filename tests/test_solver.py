"""l-BFGS with the regular Wolfe line search (fwi2d_lbfgs), on the CPU through Python
objectives: the reference's own solver tests (nlsolver.cpp:697-740: Rosenbrock from
(-1.2, 1) to its minimum (1, 1), the paraboloid from (-1, 2) to 0), a 40-dimensional
SPD quadratic solved to its closed-form minimiser with every memory size, monotone
functional histories (each accepted step satisfies sufficient decrease), the diagonal
inverse-Hessian guess, and the reference's argument checks."""
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "fwi2d_amd", "libfwi2d_hip.so")),
                                reason="library not built")


def rosenbrock(a=1.0, b=100.0):
    def f(m):
        x, y = float(m[0]), float(m[1])
        return (a - x) ** 2 + b * (y - x * x) ** 2, np.array([2 * (x - a) + 4 * b * x * (x * x - y), 2 * b * (y - x * x)])
    return f


def test_reference_rosenbrock_and_paraboloid():
    from fwi2d_amd import lbfgs
    m, hist, k = lbfgs(rosenbrock(), [-1.2, 1.0], niter=200, max_trial=20)
    print("Rosenbrock: %d iterations, minimum at %s, f %.2e" % (k, m, hist[-1]))
    assert np.hypot(m[0] - 1, m[1] - 1) < 1e-4 and np.all(np.diff(hist) < 0)
    m, hist, k = lbfgs(lambda v: (0.5 * (v[0] ** 2 + v[1] ** 2), np.array(v, np.float64)), [-1.0, 2.0], niter=50,
                       max_trial=20)
    assert np.hypot(m[0], m[1]) < 1e-6


@pytest.mark.parametrize("memory", [1, 5, 20])
def test_quadratic_minimiser(memory):
    from fwi2d_amd import lbfgs
    rng = np.random.default_rng(4)
    n = 40
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = Q @ np.diag(np.geomspace(1.0, 100.0, n)) @ Q.T
    b = rng.standard_normal(n)
    xstar = np.linalg.solve(A, b)
    fstar = -0.5 * b @ xstar

    def f(x):
        x = np.asarray(x, np.float64)
        return 0.5 * x @ A @ x - b @ x - fstar + 1.0, A @ x - b  # shifted so that the minimum is 1
    m, hist, k = lbfgs(f, np.zeros(n), niter=300, max_trial=20, memory=memory)
    err = np.linalg.norm(m - xstar) / np.linalg.norm(xstar)
    print("memory %d: %d iterations, rel err %.2e" % (memory, k, err))
    assert err < 1e-3 and np.all(np.diff(hist) <= 0)


def test_diagonal_inverse_hessian_guess():
    from fwi2d_amd import lbfgs
    d = np.geomspace(1.0, 1e4, 30)
    b = np.ones(30)

    def f(x):
        x = np.asarray(x, np.float64)
        return 0.5 * np.sum(d * x * x) - b @ x + 1.0 + 0.5 * np.sum(b * b / d), d * x - b
    _, h1, k1 = lbfgs(f, np.zeros(30), niter=200, max_trial=20, threshold=1e-12)
    _, h2, k2 = lbfgs(f, np.zeros(30), niter=200, max_trial=20, threshold=1e-12, H0=1.0 / d)
    print("iterations without / with the diagonal guess: %d / %d; f %.6e / %.6e" % (k1, k2, h1[-1], h2[-1]))
    assert k2 < k1 and h2[-1] <= 1.0 + 1e-6


def test_argument_checks():
    from fwi2d_amd import Fwi2dError, lbfgs
    with pytest.raises(Fwi2dError, match="0<c1<c2<1"):
        lbfgs(rosenbrock(), [0.0, 0.0], c1=0.5, c2=0.4)
    with pytest.raises(Fwi2dError, match="definite positive"):
        lbfgs(rosenbrock(), [0.0, 0.0], H0=[1.0, 0.0])

    def bad(m):
        raise ValueError("no")
    with pytest.raises(Fwi2dError, match="objective evaluation failed"):
        lbfgs(bad, [0.0, 0.0])
