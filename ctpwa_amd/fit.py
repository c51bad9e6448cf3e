"""Fit driver over ``analysis``: bounded L-BFGS, Hessian errors, multi-start.

Follows the reference's driver (example/fit.py:84-260, ``UnifiedPWAOptimizer``):

* the first coupling is fixed to 1 + 0i (parameters 0 and n, the global phase and scale
  that the likelihood does not see), couplings are clamped to +-v_max and resonance
  parameters to their bounds before every evaluation;
* projected gradient: at an active bound the component whose descent direction (-g)
  points out of the box is zeroed, so L-BFGS does not stop on a wall with a non-zero
  gradient.  The reference zeroes the opposite components (g < 0 at a lower bound,
  g > 0 at an upper one, example/fit.py:137-157), i.e. those whose descent direction
  points back into the box, which can keep a parameter pinned on a bound; this driver
  uses the standard projection;
* ``torch.optim.LBFGS`` with the strong-Wolfe line search (max_iter 500,
  tolerance_grad 1e-8, tolerance_change 1e-10, history 100);
* errors from the inverse of the Hessian with the two fixed parameters removed.
  ``getHessian`` uses the grouped layout [Re p, Im p, theta]; the reference's error
  extraction reads it as interleaved (example/fit.py:255-258) — here the errors are
  read in the grouped layout.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np
import torch


@dataclass
class FitResult:
    params: np.ndarray                 # [2n + P] grouped layout
    nll: float
    n_evals: int
    converged: bool
    seconds: float
    hessian: Optional[np.ndarray] = None
    errors: Optional[np.ndarray] = None   # [2n + P]; 0 for the fixed parameters
    pos_def: bool = False
    unconstrained: List[int] = field(default_factory=list)   # parameters the NLL does not depend on
    nll_history: List[float] = field(default_factory=list)


class Fitter:
    def __init__(self, ana, v_max: float = 1e4, project: bool = True):
        self.ana = ana
        self.n = ana.getNVector()
        self.P = ana.getNFreeTheta()
        self.np = 2 * self.n + self.P
        self.v_max = float(v_max)
        self.project = project
        info = ana.getFreeResParams()
        self.lower = info[1].to(dtype=torch.float64, device=ana.device)
        self.upper = info[2].to(dtype=torch.float64, device=ana.device)
        self.n_evals = 0

    def _clamp_(self, p: torch.Tensor):
        n = self.n
        with torch.no_grad():
            p[0] = 1.0
            p[n] = 0.0
            p[1:n].clamp_(-self.v_max, self.v_max)
            p[n + 1:2 * n].clamp_(-self.v_max, self.v_max)
            if self.P:
                p[2 * n:] = torch.clamp(p[2 * n:], self.lower, self.upper)

    def loss_and_grad(self, p: torch.Tensor):
        self._clamp_(p)
        nll = self.ana.getNLL(p)
        (g,) = torch.autograd.grad(nll, p)
        self.n_evals += 1
        n = self.n
        with torch.no_grad():
            g[0] = 0.0
            g[n] = 0.0
            if self.project:
                # descent direction -g leaves the box: g > 0 at a lower bound, g < 0 at an upper one
                for lo_idx, hi_idx in ((1, n), (n + 1, 2 * n)):
                    gc, c = g[lo_idx:hi_idx], p[lo_idx:hi_idx]
                    gc[(c <= -self.v_max) & (gc > 0)] = 0.0
                    gc[(c >= self.v_max) & (gc < 0)] = 0.0
                if self.P:
                    gr, r = g[2 * n:], p[2 * n:]
                    gr[(r <= self.lower) & (gr > 0)] = 0.0
                    gr[(r >= self.upper) & (gr < 0)] = 0.0
        return nll, g

    def minimize(self, p_init, max_iter: int = 500, tolerance_grad: float = 1e-8,
                 tolerance_change: float = 1e-10, history_size: int = 100, hessian: bool = True) -> FitResult:
        t0 = time.perf_counter()
        self.n_evals = 0
        p = torch.as_tensor(np.asarray(p_init, dtype=np.float64)).to(self.ana.device).clone().requires_grad_(True)
        opt = torch.optim.LBFGS([p], lr=1.0, max_iter=max_iter, tolerance_grad=tolerance_grad,
                                tolerance_change=tolerance_change, history_size=history_size,
                                line_search_fn="strong_wolfe")
        hist: List[float] = []

        def closure():
            opt.zero_grad()
            nll, g = self.loss_and_grad(p)
            p.grad = g
            hist.append(float(nll.item()))
            return nll

        opt.step(closure)
        self._clamp_(p)
        final = p.detach().clone()
        nll, g = self.loss_and_grad(final.clone().requires_grad_(True))
        state = opt.state[opt._params[0]]
        converged = bool(state.get("n_iter", 0) < max_iter)
        res = FitResult(params=final.cpu().numpy(), nll=float(nll.item()), n_evals=self.n_evals,
                        converged=converged, seconds=0.0, nll_history=hist)
        if hessian:
            self.errors(res)
        res.seconds = time.perf_counter() - t0
        return res

    def errors(self, res: FitResult) -> FitResult:
        """Hessian at the minimum, positive-definiteness, and 1-sigma errors (grouped layout)."""
        p = torch.as_tensor(res.params, device=self.ana.device)
        H = self.ana.getHessian(p)
        keep = torch.ones(self.np, dtype=torch.bool, device=H.device)
        keep[0] = False
        keep[self.n] = False
        # parameters the likelihood does not depend on (the resonance parameters of a
        # block without an allowed wave, src/Analysis.cu:4650-4675) have an all-zero
        # Hessian row: they get no error instead of making the matrix singular
        dead = torch.all(H == 0, dim=1)
        keep &= ~dead
        res.unconstrained = [int(i) for i in torch.nonzero(dead).flatten().cpu()]
        Hr = H[keep][:, keep]
        err = np.zeros(self.np)
        try:
            ev = torch.linalg.eigvalsh(Hr)
            pos = bool(torch.all(ev > 0).item())
        except RuntimeError:
            pos = False
        if pos:
            cov = torch.linalg.inv(Hr)
            err[keep.cpu().numpy()] = torch.sqrt(torch.diag(cov)).cpu().numpy()
        res.hessian = H.cpu().numpy()
        res.errors = err
        res.pos_def = pos
        return res


def fit(ana, p_init, **kw) -> FitResult:
    """One minimisation from ``p_init`` (grouped layout) with Hessian errors."""
    return Fitter(ana).minimize(p_init, **kw)


def multistart(ana, n_starts: int, seed: int = 0, scale: float = 1.0, **kw) -> List[FitResult]:
    """Fits from random couplings (|v| ~ U[0, scale], arg ~ U[0, 2 pi)) and resonance
    parameters at their initial values; results sorted by NLL."""
    f = Fitter(ana)
    rng = np.random.default_rng(seed)
    init = ana.getFreeResParams()[0].numpy()
    out = []
    for _ in range(n_starts):
        mag = rng.uniform(0, scale, f.n)
        ph = rng.uniform(0, 2 * np.pi, f.n)
        p0 = np.concatenate([mag * np.cos(ph), mag * np.sin(ph), init])
        out.append(f.minimize(p0, **kw))
    return sorted(out, key=lambda r: r.nll)
