"""The service's minimiser (ctpwa_amd.serve.lbfgs, a generator that yields points and
receives (f, g)) on problems with known minima (CPU)."""
import numpy as np

from ctpwa_amd.serve import lbfgs


def _drive(gen, fg):
    x = next(gen)
    try:
        while True:
            x = gen.send(fg(x))
    except StopIteration as s:
        return s.value


def test_bounded_quadratic_with_fixed_parameter():
    rng = np.random.default_rng(0)
    n = 8
    B = rng.standard_normal((n, n))
    A = B @ B.T + n * np.eye(n)
    mu = rng.uniform(-1, 1, n)
    lo = np.full(n, -10.0)
    hi = np.full(n, 10.0)
    lo[0] = hi[0] = 0.3                       # fixed
    hi[5] = mu[5] - 0.5                       # active upper bound
    fg = lambda x: (0.5 * (x - mu) @ A @ (x - mu), A @ (x - mu))
    out = _drive(lbfgs(np.zeros(n), lo, hi, pgtol=1e-10, ftol=0.0, max_evals=2000), fg)
    assert out.converged
    assert out.x[0] == 0.3 and out.x[5] == hi[5]
    # KKT: zero gradient on the free, inactive coordinates
    g = A @ (out.x - mu)
    free = np.ones(n, bool)
    free[[0, 5]] = False
    assert np.max(np.abs(g[free])) < 1e-6         # f no longer resolves smaller steps
    assert g[5] < 0                            # the bound is active: descent points outward


def test_rosenbrock():
    def fg(x):
        f = 100 * (x[1] - x[0] ** 2) ** 2 + (1 - x[0]) ** 2
        g = np.array([-400 * x[0] * (x[1] - x[0] ** 2) - 2 * (1 - x[0]), 200 * (x[1] - x[0] ** 2)])
        return f, g
    out = _drive(lbfgs(np.array([-1.2, 1.0]), np.full(2, -5.0), np.full(2, 5.0), pgtol=1e-9, ftol=0.0,
                       max_evals=5000), fg)
    assert np.allclose(out.x, [1.0, 1.0], atol=1e-6)
