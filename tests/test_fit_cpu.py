"""Fit driver (ctpwa_amd.fit) on a likelihood with a known minimum and Hessian (CPU).

The stand-in analysis has NLL = 1/2 (p - mu)^T A (p - mu) in the grouped layout
[Re p (n), Im p (n), theta (P)], with mu_0 = 1 and mu_n = 0 so that the fixed first
coupling sits at the minimum, one resonance parameter the NLL does not depend on, and
bounds on the others.  Checks the minimum, the errors sqrt(diag(inv(A_reduced))) read in
the grouped layout, the fixed and unconstrained parameters, and the projected gradient
at an active bound.
"""
import numpy as np
import torch

from ctpwa_amd.fit import Fitter


class QuadraticAnalysis:
    def __init__(self, n, P, mu, A, lo, hi):
        self.n, self.P = n, P
        self.mu = torch.tensor(mu)
        self.A = torch.tensor(A)
        self.lo, self.hi = lo, hi
        self.device = torch.device("cpu")

    def getNVector(self):
        return self.n

    def getNFreeTheta(self):
        return self.P

    def getFreeResParams(self):
        init = np.clip(self.mu.numpy()[2 * self.n:], self.lo, self.hi)
        return torch.tensor(np.stack([init, self.lo, self.hi]))

    def getNLL(self, p):
        d = p - self.mu
        return 0.5 * d @ (self.A @ d)

    def getHessian(self, p):
        return self.A.clone()


def _problem(seed, dead=True):
    rng = np.random.default_rng(seed)
    n, P = 3, 3
    m = 2 * n + P
    B = rng.standard_normal((m, m))
    A = B @ B.T + m * np.eye(m)
    if dead:                                     # the last theta does not enter the NLL
        A[-1, :] = 0.0
        A[:, -1] = 0.0
    mu = rng.uniform(0.5, 1.5, m)
    mu[0], mu[n] = 1.0, 0.0
    lo = np.full(P, 0.0)
    hi = np.full(P, 5.0)
    return n, P, mu, A, lo, hi


def test_minimum_and_grouped_errors():
    n, P, mu, A, lo, hi = _problem(1)
    ana = QuadraticAnalysis(n, P, mu, A, lo, hi)
    start = mu + 0.3
    start[-1] = mu[-1]
    r = Fitter(ana).minimize(start, tolerance_grad=1e-12, tolerance_change=1e-14)
    assert np.allclose(r.params[:-1], mu[:-1], atol=1e-7)
    assert r.unconstrained == [2 * n + P - 1]
    assert r.pos_def
    keep = np.ones(2 * n + P, bool)
    keep[[0, n, 2 * n + P - 1]] = False
    cov = np.linalg.inv(A[np.ix_(keep, keep)])
    exp = np.zeros(2 * n + P)
    exp[keep] = np.sqrt(np.diag(cov))
    assert np.allclose(r.errors, exp, rtol=1e-12)
    assert r.errors[0] == 0.0 and r.errors[n] == 0.0


def test_bounds_and_projected_gradient():
    n, P, mu, A, lo, hi = _problem(2, dead=False)
    hi = np.array([5.0, 5.0, 0.2])               # the minimum of the last theta lies above its bound
    ana = QuadraticAnalysis(n, P, mu, A, lo, hi)
    f = Fitter(ana)
    r = f.minimize(mu.copy(), tolerance_grad=1e-12, tolerance_change=1e-14, hessian=False)
    assert r.params[-1] == 0.2
    # at the solution the projected gradient vanishes on the free coordinates
    p = torch.tensor(r.params, requires_grad=True)
    _, g = f.loss_and_grad(p)
    assert float(g.abs().max()) < 1e-6


def test_polish_after_early_stop():
    """The damped-Newton polish (example/fit.py:308-385) takes an L-BFGS run stopped after
    one iteration to the minimum of the quadratic (with a flat direction: the damped
    branch), and is kept only when it lowers the NLL."""
    n, P, mu, A, lo, hi = _problem(3)
    ana = QuadraticAnalysis(n, P, mu, A, lo, hi)
    start = mu + 0.4
    start[-1] = mu[-1]
    raw = Fitter(ana).minimize(start, max_iter=1, hessian=False, polish=False)
    r = Fitter(ana).minimize(start, max_iter=1, hessian=False)
    assert raw.nll > 1e-3 and raw.polish_steps == 0
    assert r.polish_steps >= 1
    assert np.allclose(r.params[:-1], mu[:-1], atol=1e-6)
    assert r.nll < 1e-6          # steps gaining less than the acceptance tolerance (1e-6) stop it
    assert r.params[0] == 1.0 and r.params[n] == 0.0


def test_polish_keeps_the_box():
    """Steps are clamped into the bounds: with the minimum of the last theta above its
    upper bound the polish ends on the bound, no worse than L-BFGS alone."""
    n, P, mu, A, lo, hi = _problem(2, dead=False)
    hi = np.array([5.0, 5.0, 0.2])
    ana = QuadraticAnalysis(n, P, mu, A, lo, hi)
    f = Fitter(ana)
    p2, nll2, steps = f.polish(mu.copy())
    assert steps >= 1 and float(p2[-1]) <= 0.2
    r0 = Fitter(ana).minimize(mu.copy(), tolerance_grad=1e-12, tolerance_change=1e-14, hessian=False, polish=False)
    r1 = Fitter(ana).minimize(mu.copy(), tolerance_grad=1e-12, tolerance_change=1e-14, hessian=False)
    assert r1.nll <= r0.nll and r1.params[-1] <= 0.2


def test_multistart_checkpoint_and_warm_start(tmp_path):
    """multistart saves the best parameters after every run that reaches or lowers the best
    NLL (example/fit.py:600-650) and a warm start gives run 0 its couplings only."""
    from ctpwa_amd.fit import load_checkpoint, multistart
    n, P, mu, A, lo, hi = _problem(4)
    ana = QuadraticAnalysis(n, P, mu, A, lo, hi)
    cp = str(tmp_path / "best_params.pt")
    rs = multistart(ana, 3, seed=1, checkpoint=cp, hessian=False)
    saved = load_checkpoint(cp)
    assert saved.dtype == torch.float64 and saved.shape == (2 * n + P,)
    assert np.array_equal(saved.numpy(), rs[0].params)
    # run 0 from the saved couplings: one L-BFGS iteration and the polish reach the minimum
    w = multistart(ana, 1, seed=9, warm_start=cp, max_iter=1, hessian=False)
    assert w[0].nll <= rs[0].nll + 1e-6


def test_converged_flag_and_dead_rows_in_polish():
    """converged is False when L-BFGS stops on its iteration budget and no polish runs; the
    polish leaves the all-zero Hessian row (a parameter the NLL does not depend on) out of
    its subspace, so it takes the positive-definite Newton exit at once instead of damping
    until lam passes 1e10, and the fit then counts as converged."""
    n, P, mu, A, lo, hi = _problem(5)
    ana = QuadraticAnalysis(n, P, mu, A, lo, hi)
    start = mu + 0.4
    start[-1] = mu[-1] + 0.1
    raw = Fitter(ana).minimize(start, max_iter=1, hessian=False, polish=False)
    assert not raw.converged
    f = Fitter(ana)
    p2, nll2, steps, newton_exit = f.polish(start.copy(), with_exit=True)
    assert newton_exit and steps == 1
    assert float(p2[-1]) == start[-1]           # the dead coordinate is left unchanged
    assert np.allclose(p2.numpy()[:-1], mu[:-1], atol=1e-9)
    r = Fitter(ana).minimize(start, max_iter=1, hessian=False)
    assert r.converged and r.nll < 1e-9
