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


def _quad(seed, n=8):
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((n, n))
    A = B @ B.T + n * np.eye(n)
    mu = rng.uniform(-1, 1, n)
    lo, hi = np.full(n, -10.0), np.full(n, 10.0)
    lo[0] = hi[0] = 0.3
    hi[5] = mu[5] - 0.5
    return (lambda x: (0.5 * (x - mu) @ A @ (x - mu), A @ (x - mu))), lo, hi, mu, A


def _run_batch(problems, cap):
    from ctpwa_amd.serve import BatchLBFGS
    opt = BatchLBFGS(cap, problems[0][1].shape[0], pgtol=1e-10, ftol=0.0)
    out, active, queue = {}, {}, list(range(len(problems)))
    free = list(range(cap))
    while queue or active:
        while queue and free:
            j = queue.pop(0)
            s = free.pop(0)
            fg, lo, hi = problems[j][:3]
            opt.start(s, np.zeros_like(lo), lo, hi, max_evals=2000)
            active[s] = j
        slots = sorted(active)
        fgs = [problems[active[s]][0](opt.xt[s].copy()) for s in slots]
        done = opt.step(slots, [v[0] for v in fgs], [v[1] for v in fgs])
        for s, o in done.items():
            out[active.pop(s)] = o
            free.append(s)
    return out


def test_batch_lbfgs_solves_and_is_independent_of_the_batch():
    probs = [_quad(s) for s in range(5)]
    mix = _run_batch(probs, 3)
    for j, (fg, lo, hi, mu, A) in enumerate(probs):
        o = mix[j]
        assert o.converged
        g = A @ (o.x - mu)
        free = np.ones(len(mu), bool)
        free[[0, 5]] = False
        assert np.max(np.abs(g[free])) < 1e-6
        alone = _run_batch([probs[j]], 1)[0]
        assert alone.n_evals == o.n_evals and alone.f == o.f and np.array_equal(alone.x, o.x)
