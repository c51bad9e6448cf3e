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


def _run_newton(problems, cap, hess, x0=None):
    """Drives BatchNewton over problems (fg, lo, hi, ...) with Hessian function hess[j]."""
    from ctpwa_amd.serve import BatchNewton
    n = problems[0][1].shape[0]
    opt = BatchNewton(cap, n, edm_tol=1e-12)
    out, active, queue = {}, {}, list(range(len(problems)))
    free = list(range(cap))
    while queue or active:
        while queue and free:
            j = queue.pop(0)
            s = free.pop(0)
            fg, lo, hi = problems[j][:3]
            opt.start(s, np.zeros_like(lo) if x0 is None else x0[j], lo, hi, max_evals=200)
            active[s] = j
        slots = sorted(active)
        fgs = [problems[active[s]][0](opt.xt[s].copy()) for s in slots]
        done, need = opt.step(slots, [v[0] for v in fgs], [v[1] for v in fgs])
        if need:
            done.update(opt.set_hessians(need, np.stack([hess[active[s]](opt.x[s]) for s in need])))
        for s, o in done.items():
            out[active.pop(s)] = o
            free.append(s)
    return out


def test_batch_newton_bounded_quadratics():
    """Exact Hessians of quadratics: the fixed coordinate stays, the active bound is
    held, the free coordinates reach zero gradient in a few steps, the EDM is tiny, and
    each fit's result does not depend on the batch it ran in."""
    probs = [_quad(s) for s in range(5)]
    hess = [(lambda x, A=p[4]: A) for p in probs]
    mix = _run_newton(probs, 3, hess)
    for j, (fg, lo, hi, mu, A) in enumerate(probs):
        o = mix[j]
        assert o.converged and o.edm < 1e-12
        assert o.x[0] == 0.3 and o.x[5] == hi[5]
        g = A @ (o.x - mu)
        free = np.ones(len(mu), bool)
        free[[0, 5]] = False
        assert np.max(np.abs(g[free])) < 1e-8
        assert o.n_hess <= 4
        # errors: sqrt(diag) of the inverse Hessian on the free, inactive coordinates
        Ai = np.linalg.inv(A[np.ix_(free, free)])
        assert np.allclose(o.errors[free], np.sqrt(np.diag(Ai)), rtol=1e-10)
        assert o.errors[0] == 0.0
        alone = _run_newton([probs[j]], 1, [hess[j]])[0]
        assert alone.n_evals == o.n_evals and alone.f == o.f and np.array_equal(alone.x, o.x)


def test_batch_newton_indefinite_start():
    """A non-convex start (indefinite Hessian) is damped until positive definite and
    still reaches the minimum: f = (x0^2 - 1)^2 + x1^2 from x0 = 0.1, where f'' < 0."""
    def fg(x):
        return (x[0] ** 2 - 1) ** 2 + x[1] ** 2, np.array([4 * x[0] * (x[0] ** 2 - 1), 2 * x[1]])
    def hs(x):
        return np.diag([12 * x[0] ** 2 - 4, 2.0])
    lo, hi = np.full(2, -5.0), np.full(2, 5.0)
    o = _run_newton([(fg, lo, hi)], 1, [hs], x0=[np.array([0.1, 0.7])])[0]
    assert o.converged
    assert np.allclose(o.x, [1.0, 0.0], atol=1e-7)


def test_batch_newton_toy_fit_on_the_oracle(oracle_lib):
    """A toy fit of the three-pion model driven by the oracle's NLL, gradient and Hessian:
    Newton converges within a few Hessians, its point is a minimum (projected gradient
    ~0, and L-BFGS started there finds nothing lower), and it is lower than where L-BFGS
    from the same start stands after 400 evaluations."""
    from helpers import plan_of
    from ctpwa_amd import workloads as W
    from ctpwa_amd.serve import toy_ensemble
    plan = plan_of("jpsi_3pi")

    def intensity(plan_text, ev, params):
        return oracle_lib.Oracle(plan_text, 3, ev, ev[:1], None).intensity(params, 0)
    s = W.generate_samples("jpsi_3pi", intensity, n_data=3000, n_phsp=15000, n_bkg=0, seed=11)
    job = toy_ensemble(plan, s["phsp"], s["data"], 1, seed=5, n_min=3000, n_max=3000,
                       x_nominal=W.nominal_params(plan, 1))[0]
    ev = job.samples()
    O = oracle_lib.Oracle(plan.to_text(), 3, ev["phsp"], ev["data"], None)
    fg = lambda x: O.nll_grad(x)
    hs = lambda x: O.hessian(x)[2]
    o = _run_newton([(fg, job.lower, job.upper)], 1, [hs], x0=[job.x0])[0]
    assert o.converged and o.n_hess <= 12
    _, g = O.nll_grad(o.x)
    g[job.lower == job.upper] = 0.0
    assert np.max(np.abs(g)) < 1e-3
    polish = _drive(lbfgs(o.x, job.lower, job.upper, max_evals=60), fg)
    assert polish.f >= o.f - 1e-7
    slow = _drive(lbfgs(job.x0, job.lower, job.upper, max_evals=400), fg)
    assert o.f <= slow.f + 1e-9
