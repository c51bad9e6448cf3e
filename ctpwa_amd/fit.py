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
* after L-BFGS, a damped-Newton polish on the exact Hessian (example/fit.py:308-385,
  on unless ``polish=False``, as the reference's FIT_POLISH=1 default), kept when it
  lowers the NLL;
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
    polish_steps: int = 0              # damped-Newton steps taken after L-BFGS (0: none)


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
                 tolerance_change: float = 1e-10, history_size: int = 100, hessian: bool = True,
                 polish: bool = True) -> FitResult:
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
        max_eval = opt.param_groups[0]["max_eval"]
        # L-BFGS converged when it stopped on its own tolerances, not on its iteration or
        # evaluation budget
        converged = bool(state.get("n_iter", 0) < max_iter and state.get("func_evals", 0) < max_eval)
        steps = 0
        if polish:
            p2, nll2, steps, newton_exit = self.polish(final, with_exit=True)
            moved = nll2 < float(nll.item())
            if moved:
                final, nll = p2, torch.tensor(nll2)
            # after the polish: converged when it ended on the positive-definite Newton exit,
            # or when L-BFGS had converged and the polish left the point unchanged
            converged = bool(newton_exit or (converged and not moved))
        res = FitResult(params=final.cpu().numpy(), nll=float(nll.item()), n_evals=self.n_evals,
                        converged=converged, seconds=0.0, nll_history=hist, polish_steps=steps)
        if hessian:
            self.errors(res)
        res.seconds = time.perf_counter() - t0
        return res

    def polish(self, p_init, max_steps: int = 40, tol: float = 1e-6, lam0: float = 1.0, with_exit: bool = False):
        """Damped Newton on the exact Hessian (example/fit.py:308-385).  In the subspace
        without the two fixed parameters, solve (H + lam I) d = -g with lam at least
        -lambda_min(H) + 1e-6, so that d descends where H is not positive definite; the
        step is clamped into the box; an accepted step (NLL lower by more than tol) scales
        lam by 0.3 (not below 1e-6), a rejected one by 10, and lam above 1e10 ends it.  Once
        H is positive definite, one undamped Newton step (kept if it lowers the NLL) ends
        it.  Parameters with an all-zero Hessian row (the likelihood does not depend on
        them, as errors() treats them) stay out of the subspace and unchanged.  Returns
        (params, nll, steps), and with ``with_exit`` also whether the positive-definite
        Newton exit was reached."""
        dev = self.ana.device
        n = self.n
        best = torch.as_tensor(np.asarray(p_init.cpu() if torch.is_tensor(p_init) else p_init,
                                          dtype=np.float64)).to(dev).clone()
        self._clamp_(best)
        mask = torch.ones(self.np, dtype=torch.bool, device=dev)
        mask[0] = False
        mask[n] = False
        H_full = self.ana.getHessian(best)
        mask &= ~torch.all(H_full == 0, dim=1)
        nf = int(mask.sum().item())
        newton_exit = False

        def nll_at(q):
            self.n_evals += 1
            return float(self.ana.getNLL(q).item())

        nll_best = nll_at(best)
        lam = lam0
        steps = 0
        for step in range(max_steps):
            steps = step + 1
            if H_full is None:
                H_full = self.ana.getHessian(best)
            H = H_full[mask][:, mask]
            H_full = None
            eig_min = float(torch.linalg.eigvalsh(H)[0].item())
            q = best.clone().requires_grad_(True)
            self.n_evals += 1
            (g,) = torch.autograd.grad(self.ana.getNLL(q), q)
            gm = g[mask]
            if eig_min > 1e-8:
                cand = best.clone()
                cand[mask] = best[mask] + torch.linalg.solve(H, -gm)
                self._clamp_(cand)
                nc = nll_at(cand)
                if nc < nll_best - tol:
                    best, nll_best = cand, nc
                newton_exit = True
                break
            lam = max(lam, -eig_min + 1e-6)
            Hd = H + lam * torch.eye(nf, dtype=torch.float64, device=dev)
            try:
                d = torch.linalg.solve(Hd, -gm)
            except RuntimeError:
                d = torch.linalg.lstsq(Hd, -gm.unsqueeze(1)).solution.squeeze(1)
            cand = best.clone()
            cand[mask] = best[mask] + d
            self._clamp_(cand)
            nc = nll_at(cand)
            if nc < nll_best - tol:
                best, nll_best = cand, nc
                lam = max(lam * 0.3, 1e-6)
            else:
                lam *= 10.0
                if lam > 1e10:
                    break
        if with_exit:
            return best.detach(), nll_best, steps, newton_exit
        return best.detach(), nll_best, steps

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


def multistart(ana, n_starts: int, seed: int = 0, scale: float = 1.0, checkpoint: Optional[str] = None,
               warm_start=None, **kw) -> List[FitResult]:
    """Fits from random couplings (|v| ~ U[0, scale], arg ~ U[0, 2 pi)) and resonance
    parameters at their initial values; results sorted by NLL.  As the reference's
    multi-start run (example/fit.py:600-650): ``warm_start`` (parameters, or the path of a
    checkpoint) gives run 0 its couplings, the resonance parameters staying at their
    initial values; with ``checkpoint`` the best parameters so far are saved there (a CPU
    float64 tensor, ``torch.save``) after every run that reaches or lowers the best NLL."""
    f = Fitter(ana)
    rng = np.random.default_rng(seed)
    init = ana.getFreeResParams()[0].numpy()
    if isinstance(warm_start, str):
        warm_start = load_checkpoint(warm_start)
    out = []
    best = float("inf")
    for i in range(n_starts):
        mag = rng.uniform(0, scale, f.n)
        ph = rng.uniform(0, 2 * np.pi, f.n)
        p0 = np.concatenate([mag * np.cos(ph), mag * np.sin(ph), init])
        if i == 0 and warm_start is not None:
            w = np.asarray(warm_start.cpu() if torch.is_tensor(warm_start) else warm_start, dtype=np.float64)
            p0[:2 * f.n] = w[:2 * f.n]
        r = f.minimize(p0, **kw)
        out.append(r)
        if checkpoint and r.nll <= best:
            best = r.nll
            torch.save(torch.as_tensor(r.params, dtype=torch.float64), checkpoint)
    return sorted(out, key=lambda r: r.nll)


def load_checkpoint(path: str) -> torch.Tensor:
    """Parameters saved by ``multistart(checkpoint=...)`` (or the reference's
    best_params.pt), read with the loader that executes nothing from the file."""
    return torch.load(path, weights_only=True).to(torch.float64)
