"""Resident fit service with continuous batching of NLL+gradient requests.

Many fits of one model (multi-start minimisations, toy-MC ensembles with uneven event
counts) share the GPU.  All active fits advance in lock step:

    round:  gather the pending point of every active fit
            -> one batched evaluation (ctpwa_batch_nll_grad: batched prep, one grouped
               per-event launch whose workgroups are shared out among the fits, batched
               reduction) on the service's stream
            -> one step of the minimiser for all fits at once; finished fits retire and
               free their slot and HBM
            -> (Newton) the exact Hessian (ctpwa_hessian) of each fit that moved
    meanwhile a worker thread admits arrived fits in order: it waits for HBM
    (ctpwa_footprint_bytes against the budget) and builds each fit's handle (upload,
    SL tensors) while the rounds go on; ready fits take free slots between rounds.

Each fit's events stay resident on the device for its lifetime.  A fit's results are
bit-identical whether it runs alone or in a batch (the batched evaluation reproduces the
single one exactly, and the minimiser reduces every fit's row on its own), so a fit's
trajectory does not depend on what else is running.

The reference has no service; its fit driver (example/fit.py:84-260) runs one fit per
process with torch's L-BFGS and takes parameter errors from getHessian at the end.  The
service's default minimiser, BatchNewton, uses that exact Hessian at every iterate: a
bounded, Levenberg-Marquardt-damped Newton method that stops on the estimated distance to
the minimum EDM = g^T H^-1 g / 2 (the MINUIT criterion) and returns the errors
sqrt(diag H^-1) of its last step.  Partial-wave likelihoods have long, flat valleys
(broad overlapping resonances); quasi-Newton steps crawl along them, so an L-BFGS fit of a
few thousand events is still lowering the NLL after hundreds of evaluations (scipy's
L-BFGS-B behaves the same), while Newton steps take a handful.  Models whose Hessian is not
built (Custom lineshapes) use BatchLBFGS: a bounded L-BFGS (two-loop recursion, projected
gradient, Armijo backtracking) kept on the host in float64 numpy; ``lbfgs`` is the same
algorithm as a generator for a single fit.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import Callable, Dict, Generator, List, Optional, Tuple

import numpy as np
import torch

from . import _native


# ---------------------------------------------------------------------------
# minimiser as a generator
# ---------------------------------------------------------------------------
@dataclass
class Outcome:
    x: np.ndarray
    f: float
    n_evals: int
    n_iter: int
    converged: bool
    n_hess: int = 0
    edm: float = float("nan")
    errors: Optional[np.ndarray] = None      # sqrt(diag H^-1); 0 for fixed, nan for flat parameters


def lbfgs(x0: np.ndarray, lower: np.ndarray, upper: np.ndarray, m: int = 10, max_evals: int = 400,
          pgtol: float = 1e-5, ftol: float = 1e-11) -> Generator[np.ndarray, Tuple[float, np.ndarray], Outcome]:
    """Bounded L-BFGS.  Yields points, receives (f, g); returns an Outcome.

    Parameters with lower == upper are fixed.  Convergence: the largest projected-gradient
    component below ``pgtol * max(1, |f|)`` or a relative decrease of f below ``ftol``.
    """
    lower = np.asarray(lower, dtype=np.float64)
    upper = np.asarray(upper, dtype=np.float64)
    x = np.clip(np.asarray(x0, dtype=np.float64), lower, upper)
    free = lower < upper
    f, g = yield x.copy()
    n_evals = 1
    S: List[np.ndarray] = []
    Y: List[np.ndarray] = []
    it = 0

    def projected(x, g):
        pg = np.where(free, g, 0.0)
        pg = np.where((x <= lower) & (pg > 0), 0.0, pg)
        pg = np.where((x >= upper) & (pg < 0), 0.0, pg)
        return pg

    restarted = False
    while n_evals < max_evals:
        pg = projected(x, g)
        if np.max(np.abs(pg)) <= pgtol * max(1.0, abs(f)):
            return Outcome(x, f, n_evals, it, True)
        # two-loop recursion restricted to the inactive coordinates (free, and not held
        # at a bound by a gradient that points out of the box)
        act = pg == 0.0
        q = pg.copy()
        alpha = []
        for s_, y_ in zip(reversed(S), reversed(Y)):
            s, y = np.where(act, 0.0, s_), np.where(act, 0.0, y_)
            sy = y @ s
            if sy <= 0:
                alpha.append(0.0)
                continue
            a = s @ q / sy
            alpha.append(a)
            q -= a * y
        if S:
            s, y = np.where(act, 0.0, S[-1]), np.where(act, 0.0, Y[-1])
            yy = y @ y
            q *= (s @ y) / yy if yy > 0 and s @ y > 0 else 1.0
        else:
            q *= 1.0 / max(1.0, np.linalg.norm(pg))
        for (s_, y_), a in zip(zip(S, Y), reversed(alpha)):
            s, y = np.where(act, 0.0, s_), np.where(act, 0.0, y_)
            sy = y @ s
            if sy <= 0:
                continue
            b = y @ q / sy
            q += (a - b) * s
        d = -np.where(act, 0.0, q)
        if pg @ d >= 0:                      # not a descent direction: steepest descent
            S.clear()
            Y.clear()
            d = -pg / max(1.0, np.linalg.norm(pg))
        t = 1.0
        accepted = False
        while n_evals < max_evals:
            xn = np.clip(x + t * d, lower, upper)
            fn, gn = yield xn.copy()
            n_evals += 1
            if np.isfinite(fn) and fn <= f + 1e-4 * (pg @ (xn - x)):
                accepted = True
                break
            t *= 0.5
            if t < 1e-12:
                break
        if not accepted:
            if S and not restarted:          # retry once from steepest descent
                S.clear()
                Y.clear()
                restarted = True
                continue
            return Outcome(x, f, n_evals, it, False)
        restarted = False
        s, y = xn - x, gn - g
        if s @ y > 1e-12 * np.linalg.norm(s) * np.linalg.norm(y):
            S.append(s)
            Y.append(y)
            if len(S) > m:
                S.pop(0)
                Y.pop(0)
        df = f - fn
        x, f, g = xn, fn, gn
        it += 1
        if 0 <= df <= ftol * max(1.0, abs(f)):
            pgn = projected(x, g)
            return Outcome(x, f, n_evals, it, bool(np.max(np.abs(pgn)) <= 1e3 * pgtol * max(1.0, abs(f))))
    return Outcome(x, f, n_evals, it, False)


# ---------------------------------------------------------------------------
# the same minimiser for many fits at once (what the service runs)
# ---------------------------------------------------------------------------
def _rowdot(a, b):
    """Row-wise dot products; each row is reduced on its own, so a fit's numbers do not
    depend on which other fits share the arrays."""
    return np.add.reduce(a * b, axis=-1)


class BatchLBFGS:
    """Bounded L-BFGS (the algorithm of :func:`lbfgs`) advanced for many fits per call,
    with numpy operations over the fits' rows.  ``start`` places a fit in a slot;
    ``step(slots, f, g)`` consumes the results of the slots' pending points and returns
    the slots that finished (with their Outcome); ``xt[slot]`` is each remaining slot's
    next point."""

    def __init__(self, cap: int, n: int, m: int = 10, pgtol: float = 1e-5, ftol: float = 1e-11):
        self.cap, self.n, self.m, self.pgtol, self.ftol = cap, n, m, pgtol, ftol
        z = lambda *sh: np.zeros(sh)
        self.x, self.g, self.xt, self.d = z(cap, n), z(cap, n), z(cap, n), z(cap, n)
        self.lo, self.hi = z(cap, n), z(cap, n)
        self.f, self.t = z(cap), z(cap)
        self.S, self.Y = z(cap, m, n), z(cap, m, n)
        self.hl = np.zeros(cap, dtype=np.int64)
        self.phase = np.zeros(cap, dtype=np.int64)        # 0: first point pending, 1: trial pending
        self.n_evals = np.zeros(cap, dtype=np.int64)
        self.n_iter = np.zeros(cap, dtype=np.int64)
        self.max_evals = np.zeros(cap, dtype=np.int64)
        self.restarted = np.zeros(cap, dtype=bool)

    def start(self, slot: int, x0, lower, upper, max_evals: int = 400) -> np.ndarray:
        self.lo[slot], self.hi[slot] = lower, upper
        self.xt[slot] = np.clip(np.asarray(x0, dtype=np.float64), lower, upper)
        self.phase[slot] = 0
        self.hl[slot] = 0
        self.n_evals[slot] = 0
        self.n_iter[slot] = 0
        self.max_evals[slot] = max_evals
        self.restarted[slot] = False
        return self.xt[slot]

    def _projected(self, idx):
        x, g, lo, hi = self.x[idx], self.g[idx], self.lo[idx], self.hi[idx]
        pg = np.where(lo < hi, g, 0.0)
        pg = np.where((x <= lo) & (pg > 0), 0.0, pg)
        pg = np.where((x >= hi) & (pg < 0), 0.0, pg)
        return pg

    def _directions(self, idx, finished):
        """New search directions for the slots idx (x, f, g current); slots that converge
        are added to ``finished``."""
        if len(idx) == 0:
            return
        pg = self._projected(idx)
        scale = np.maximum(1.0, np.abs(self.f[idx]))
        conv = np.max(np.abs(pg), axis=1) <= self.pgtol * scale
        for k in np.nonzero(conv)[0]:
            finished[int(idx[k])] = True
        keep = ~conv
        idx, pg = idx[keep], pg[keep]
        if len(idx) == 0:
            return
        act = pg == 0.0
        q = pg.copy()
        hl = self.hl[idx]
        alpha = np.zeros((len(idx), self.m))
        for k in range(self.m - 1, -1, -1):
            valid = k < hl
            s_ = np.where(act, 0.0, self.S[idx, k])
            y_ = np.where(act, 0.0, self.Y[idx, k])
            sy = _rowdot(y_, s_)
            use = valid & (sy > 0)
            a = np.where(use, _rowdot(s_, q) / np.where(use, sy, 1.0), 0.0)
            alpha[:, k] = a
            q = q - a[:, None] * y_
        last = np.maximum(hl - 1, 0)
        s_ = np.where(act, 0.0, self.S[idx, last])
        y_ = np.where(act, 0.0, self.Y[idx, last])
        sy, yy = _rowdot(s_, y_), _rowdot(y_, y_)
        gamma = np.where((hl > 0) & (yy > 0) & (sy > 0), sy / np.where(yy > 0, yy, 1.0), 1.0)
        first = 1.0 / np.maximum(1.0, np.sqrt(_rowdot(pg, pg)))
        q = q * np.where(hl > 0, gamma, first)[:, None]
        for k in range(self.m):
            valid = k < hl
            s_ = np.where(act, 0.0, self.S[idx, k])
            y_ = np.where(act, 0.0, self.Y[idx, k])
            sy = _rowdot(y_, s_)
            use = valid & (sy > 0)
            b = np.where(use, _rowdot(y_, q) / np.where(use, sy, 1.0), 0.0)
            q = q + np.where(use, alpha[:, k] - b, 0.0)[:, None] * s_
        d = -np.where(act, 0.0, q)
        bad = _rowdot(pg, d) >= 0
        if np.any(bad):
            self.hl[idx[bad]] = 0
            d[bad] = -pg[bad] / np.maximum(1.0, np.sqrt(_rowdot(pg[bad], pg[bad])))[:, None]
        self.d[idx] = d
        self.t[idx] = 1.0
        self.xt[idx] = np.clip(self.x[idx] + d, self.lo[idx], self.hi[idx])

    def step(self, slots, f, g) -> Dict[int, Outcome]:
        slots = np.asarray(slots, dtype=np.int64)
        f = np.asarray(f, dtype=np.float64)
        g = np.asarray(g, dtype=np.float64)
        self.n_evals[slots] += 1
        finished: Dict[int, bool] = {}
        need_dir = []
        # first evaluations
        ini = self.phase[slots] == 0
        si = slots[ini]
        self.x[si], self.f[si], self.g[si] = self.xt[si], f[ini], g[ini]
        self.phase[si] = 1
        need_dir.append(si)
        # trial evaluations
        tr = ~ini
        st, ft, gt = slots[tr], f[tr], g[tr]
        if len(st):
            pg = self._projected(st)
            ok = np.isfinite(ft) & (ft <= self.f[st] + 1e-4 * _rowdot(pg, self.xt[st] - self.x[st]))
            # accepted: history update, move, convergence on the decrease
            sa = st[ok]
            if len(sa):
                s_ = self.xt[sa] - self.x[sa]
                y_ = gt[ok] - self.g[sa]
                good = _rowdot(s_, y_) > 1e-12 * np.sqrt(_rowdot(s_, s_)) * np.sqrt(_rowdot(y_, y_))
                for k in np.nonzero(good)[0]:
                    sl = int(sa[k])
                    h = int(self.hl[sl])
                    if h == self.m:
                        self.S[sl, :-1] = self.S[sl, 1:]
                        self.Y[sl, :-1] = self.Y[sl, 1:]
                        h -= 1
                    self.S[sl, h] = s_[k]
                    self.Y[sl, h] = y_[k]
                    self.hl[sl] = h + 1
                df = self.f[sa] - ft[ok]
                self.x[sa], self.f[sa], self.g[sa] = self.xt[sa], ft[ok], gt[ok]
                self.n_iter[sa] += 1
                self.restarted[sa] = False
                small = (df >= 0) & (df <= self.ftol * np.maximum(1.0, np.abs(self.f[sa])))
                for k in np.nonzero(small)[0]:
                    finished[int(sa[k])] = False
                need_dir.append(sa[~small])
            # rejected: backtrack, or restart once from steepest descent
            sr = st[~ok]
            if len(sr):
                self.t[sr] *= 0.5
                tiny = self.t[sr] < 1e-12
                back = sr[~tiny]
                self.xt[back] = np.clip(self.x[back] + self.t[back, None] * self.d[back], self.lo[back], self.hi[back])
                for sl in sr[tiny]:
                    sl = int(sl)
                    if self.hl[sl] > 0 and not self.restarted[sl]:
                        self.hl[sl] = 0
                        self.restarted[sl] = True
                        need_dir.append(np.array([sl], dtype=np.int64))
                    else:
                        finished[sl] = None
        conv: Dict[int, bool] = {}
        nd = np.concatenate(need_dir) if need_dir else np.zeros(0, dtype=np.int64)
        self._directions(nd, conv)
        out: Dict[int, Outcome] = {}
        for sl, c in conv.items():
            out[sl] = Outcome(self.x[sl].copy(), float(self.f[sl]), int(self.n_evals[sl]), int(self.n_iter[sl]), True)
        for sl, flag in finished.items():
            if flag is None:
                c = False
            else:
                pgn = self._projected(np.array([sl]))[0]
                c = bool(np.max(np.abs(pgn)) <= 1e3 * self.pgtol * max(1.0, abs(self.f[sl])))
            out[sl] = Outcome(self.x[sl].copy(), float(self.f[sl]), int(self.n_evals[sl]), int(self.n_iter[sl]), c)
        # evaluation budget
        for sl in slots:
            sl = int(sl)
            if sl not in out and self.n_evals[sl] >= self.max_evals[sl]:
                out[sl] = Outcome(self.x[sl].copy(), float(self.f[sl]), int(self.n_evals[sl]), int(self.n_iter[sl]), False)
        return out


class BatchNewton:
    """Bounded damped Newton with exact Hessians, for many fits.

    ``start`` places a fit in a slot; ``step(slots, f, g)`` consumes the results of the
    slots' pending points and returns (finished, need_hessian): the slots whose point was
    accepted (or was the first one) need the Hessian there, given back with
    ``set_hessians(slots, H)``, which computes each slot's next trial point ``xt[slot]``
    or finishes it.  Every slot is handled on its own rows, so a fit's trajectory does not
    depend on the other slots.

    Direction: coordinates that are fixed (lower == upper), held at a bound by a gradient
    pointing out of the box, or flat (zero Hessian row, e.g. a resonance with no allowed
    wave) are left out; on the rest, (H + lam diag|H|) d = -g with lam = 0 when H is
    positive definite, else the smallest lam in 1e-8 * 10^k that makes it so
    (Levenberg-Marquardt); lam shrinks tenfold after every accepted step.  Armijo
    backtracking halves the step; after four halvings lam grows tenfold.  Converged: the
    undamped H positive definite on the free coordinates and EDM = g^T H^-1 g / 2 < edm_tol
    (checked at every Hessian, whatever the damping).  The errors are sqrt(diag H^-1) there."""

    def __init__(self, cap: int, n: int, edm_tol: float = 1e-5, c1: float = 1e-4):
        self.cap, self.n, self.edm_tol, self.c1 = cap, n, edm_tol, c1
        z = lambda *sh: np.zeros(sh)
        self.x, self.g, self.xt, self.d = z(cap, n), z(cap, n), z(cap, n), z(cap, n)
        self.lo, self.hi = z(cap, n), z(cap, n)
        self.H = z(cap, n, n)
        self.f, self.t, self.lam, self.edm = z(cap), z(cap), z(cap), z(cap)
        self.phase = np.zeros(cap, dtype=np.int64)        # 0: first point pending, 1: trial pending
        self.halvings = np.zeros(cap, dtype=np.int64)
        self.n_evals = np.zeros(cap, dtype=np.int64)
        self.n_hess = np.zeros(cap, dtype=np.int64)
        self.n_iter = np.zeros(cap, dtype=np.int64)
        self.max_evals = np.zeros(cap, dtype=np.int64)

    def start(self, slot: int, x0, lower, upper, max_evals: int = 400) -> np.ndarray:
        self.lo[slot], self.hi[slot] = lower, upper
        self.xt[slot] = np.clip(np.asarray(x0, dtype=np.float64), lower, upper)
        self.phase[slot] = 0
        self.lam[slot] = 0.0
        self.edm[slot] = np.nan
        self.halvings[slot] = 0
        self.n_evals[slot] = self.n_hess[slot] = self.n_iter[slot] = 0
        self.max_evals[slot] = max_evals
        return self.xt[slot]

    def _outcome(self, sl: int, converged: bool, errors=None) -> Outcome:
        return Outcome(self.x[sl].copy(), float(self.f[sl]), int(self.n_evals[sl]), int(self.n_iter[sl]), converged,
                       int(self.n_hess[sl]), float(self.edm[sl]), errors)

    def step(self, slots, f, g) -> Tuple[Dict[int, Outcome], List[int]]:
        out: Dict[int, Outcome] = {}
        need_h: List[int] = []
        for k, sl in enumerate(np.asarray(slots, dtype=np.int64)):
            sl = int(sl)
            self.n_evals[sl] += 1
            fk, gk = float(f[k]), np.asarray(g[k], dtype=np.float64)
            if self.phase[sl] == 0:
                self.x[sl], self.f[sl], self.g[sl] = self.xt[sl], fk, gk
                need_h.append(sl)
                continue
            ok = np.isfinite(fk) and fk <= self.f[sl] + self.c1 * float(self.g[sl] @ (self.xt[sl] - self.x[sl]))
            if ok:
                self.x[sl], self.f[sl], self.g[sl] = self.xt[sl], fk, gk
                self.n_iter[sl] += 1
                self.lam[sl] *= 0.1 if self.lam[sl] > 1e-7 else 0.0
                need_h.append(sl)
            else:
                self.halvings[sl] += 1
                if self.halvings[sl] > 40:
                    out[sl] = self._outcome(sl, False)
                    continue
                if self.halvings[sl] % 4 == 0:
                    self.lam[sl] = max(10.0 * self.lam[sl], 1e-6)
                    r = self._direction(sl)
                    if r is not None:
                        out[sl] = r
                        continue
                else:
                    self.t[sl] *= 0.5
                    self.xt[sl] = np.clip(self.x[sl] + self.t[sl] * self.d[sl], self.lo[sl], self.hi[sl])
            if sl not in out and self.n_evals[sl] >= self.max_evals[sl]:
                out[sl] = self._outcome(sl, False)
        need_h = [s for s in need_h if s not in out]
        for sl in need_h:
            if self.n_evals[sl] >= self.max_evals[sl]:
                out[sl] = self._outcome(sl, False)
        return out, [s for s in need_h if s not in out]

    def set_hessians(self, slots, H) -> Dict[int, Outcome]:
        out: Dict[int, Outcome] = {}
        for k, sl in enumerate(slots):
            sl = int(sl)
            self.H[sl] = H[k]
            self.n_hess[sl] += 1
            self.phase[sl] = 1
            self.halvings[sl] = 0
            r = self._direction(sl)
            if r is not None:
                out[sl] = r
        return out

    def _direction(self, sl: int) -> Optional[Outcome]:
        x, g, H, lo, hi = self.x[sl], self.g[sl], self.H[sl], self.lo[sl], self.hi[sl]
        dg = np.abs(np.diag(H))
        held = ((x <= lo) & (g > 0)) | ((x >= hi) & (g < 0))
        flat = (dg == 0.0) & (g == 0.0)
        free = (lo < hi) & ~held & ~flat
        idx = np.nonzero(free)[0]
        if len(idx) == 0:
            self.edm[sl] = 0.0
            return self._outcome(sl, True)
        Hf, gf, Df = H[np.ix_(idx, idx)], g[idx], np.diag(np.where(dg[idx] > 0, dg[idx], 1.0))
        # convergence is judged on the undamped Hessian, whatever the current damping
        try:
            L0 = np.linalg.cholesky(Hf)
        except np.linalg.LinAlgError:
            L0 = None
        if L0 is not None:
            dn0 = -np.linalg.solve(L0.T, np.linalg.solve(L0, gf))
            self.edm[sl] = -0.5 * float(gf @ dn0)
            if self.edm[sl] < self.edm_tol:
                inv = np.linalg.solve(L0.T, np.linalg.solve(L0, np.eye(len(idx))))
                err = np.zeros(self.n)
                err[(lo < hi) & flat] = np.nan
                err[idx] = np.sqrt(np.diag(inv))
                return self._outcome(sl, True, err)
        lam = self.lam[sl]
        L = L0 if (lam == 0.0 and L0 is not None) else None
        for _ in range(40):
            if L is not None:
                break
            try:
                L = np.linalg.cholesky(Hf + lam * Df)
            except np.linalg.LinAlgError:
                lam = max(10.0 * lam, 1e-8)
        if L is None:
            return self._outcome(sl, False)
        self.lam[sl] = lam
        dn = -np.linalg.solve(L.T, np.linalg.solve(L, gf))
        d = np.zeros(self.n)
        d[idx] = dn
        self.d[sl] = d
        self.t[sl] = 1.0
        self.xt[sl] = np.clip(x + d, lo, hi)
        return None


# ---------------------------------------------------------------------------
# service
# ---------------------------------------------------------------------------
@dataclass
class FitJob:
    """One fit: its events (arrays [N, nfinal, 4] or a callable that returns them at
    admission), its starting point and bounds, and its arrival time (s after start)."""
    job_id: int
    samples: object
    x0: np.ndarray
    lower: np.ndarray
    upper: np.ndarray
    arrival: float = 0.0
    max_evals: int = 400


@dataclass
class FitRecord:
    """``status``: "ok" (the minimiser ran to its end), "rejected" (not admitted: its HBM
    footprint exceeds the service's budget, or its starting point is not finite) or
    "failed" (its events could not be loaded, its handle could not be built, or its
    likelihood or Hessian failed or was not finite at its starting point); ``error`` says
    why.  A rejected or failed fit leaves every other fit's result unchanged."""
    job_id: int
    outcome: Outcome
    n_events: int
    admitted: float
    finished: float
    arrival: float = 0.0
    status: str = "ok"
    error: str = ""


@dataclass
class ServiceStats:
    fits: List[FitRecord] = field(default_factory=list)
    latencies: List[float] = field(default_factory=list)     # per evaluation, seconds
    batch_sizes: List[int] = field(default_factory=list)
    round_seconds: List[float] = field(default_factory=list)
    hessian_seconds: List[float] = field(default_factory=list)   # per round with Hessians
    wall: float = 0.0
    admissions_seconds: float = 0.0

    def summary(self) -> Dict[str, float]:
        """Service figures over the fits that ran ("ok"); rejected and failed fits are
        counted apart."""
        lat = np.array(self.latencies) if self.latencies else np.zeros(1)
        ok = [r for r in self.fits if r.status == "ok"]
        fit_lat = [r.finished - r.arrival for r in ok]
        return {
            "fits": len(ok),
            "rejected": int(sum(r.status == "rejected" for r in self.fits)),
            "failed": int(sum(r.status == "failed" for r in self.fits)),
            "evaluations": int(len(self.latencies)),
            "wall_s": self.wall,
            "fits_per_hour": len(ok) / self.wall * 3600.0 if self.wall > 0 else 0.0,
            "evals_per_s": len(self.latencies) / self.wall if self.wall > 0 else 0.0,
            "latency_p50_ms": float(np.percentile(lat, 50) * 1e3),
            "latency_p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_batch": float(np.mean(self.batch_sizes)) if self.batch_sizes else 0.0,
            "rounds": len(self.batch_sizes),
            "admission_s": self.admissions_seconds,
            "converged": int(sum(r.outcome.converged for r in ok)),
            "hessians": int(sum(r.outcome.n_hess for r in ok)),
            "evals_per_fit_median": float(np.median([r.outcome.n_evals for r in ok])) if ok else 0.0,
            "hessian_round_p50_ms": float(np.percentile(self.hessian_seconds, 50) * 1e3) if self.hessian_seconds else 0.0,
            # arrival -> result of whole fits (admission queueing included)
            "fit_latency_p50_s": float(np.percentile(fit_lat, 50)) if ok else 0.0,
            "fit_latency_p99_s": float(np.percentile(fit_lat, 99)) if ok else 0.0,
        }


class FitService:
    """Runs fits of one model (``plan_text``) with continuous batching on one GPU."""

    def __init__(self, plan_text: str, precision="fp64", max_slots: int = 64, hbm_budget: Optional[int] = None,
                 device: int = 0, clock: Callable[[], float] = time.perf_counter, method: str = "auto",
                 edm_tol: float = 1e-5):
        """``method``: "newton" (exact Hessians, BatchNewton), "lbfgs" (BatchLBFGS), or
        "auto" (Newton)."""
        if method == "auto":
            method = "newton"
        if method not in ("newton", "lbfgs"):
            raise ValueError("method must be 'auto', 'newton' or 'lbfgs'")
        self.method = method
        self.edm_tol = edm_tol
        self.plan_text = plan_text
        self.precision = _native.precision_code(precision)
        self.max_slots = max_slots
        self.device = device
        if hbm_budget is None:
            free, _total = torch.cuda.mem_get_info(device)
            hbm_budget = int(free * 0.8)
        self.hbm_budget = hbm_budget
        self.clock = clock
        self.batch = _native.Batch(max_slots, device=device)
        self.np = None

    def footprint(self, n3) -> int:
        b = _native.footprint_bytes(self.plan_text, n3[0], n3[1], n3[2], self.precision)
        if self.method == "newton":
            b += _native.hessian_workspace_bytes(self.plan_text, n3[0], n3[1], n3[2])
        return b

    def run(self, jobs: List[FitJob]) -> ServiceStats:
        """Serve ``jobs`` until every fit has finished.

        A worker thread admits fits in arrival order: it waits for the arrival time and
        for HBM (ctpwa_footprint_bytes against the budget), then builds the fit's handle
        (upload, SL tensors) while the main loop keeps evaluating the active fits on its
        own stream.  The main loop moves ready fits into free slots between rounds."""
        import queue as _queue
        import threading
        st = ServiceStats()
        order = sorted(jobs, key=lambda j: j.arrival)
        dev = torch.device("cuda", self.device)
        stream = torch.cuda.Stream(device=dev)
        ready: "_queue.Queue" = _queue.Queue()
        cond = threading.Condition()
        mem = {"used": 0, "stop": False}
        t0 = self.clock()

        def admit():
            # a job that cannot be admitted is reported and skipped; only an error of the
            # service itself (not of one job) ends the run
            try:
                for job in order:
                    while not mem["stop"]:
                        wait = job.arrival - (self.clock() - t0)
                        if wait <= 0:
                            break
                        time.sleep(min(wait, 0.005))
                    if mem["stop"]:
                        return
                    n3 = (0, 0, 0)
                    try:
                        if not np.all(np.isfinite(np.asarray(job.x0, dtype=np.float64))):
                            ready.put(("rejected", job, n3, "the starting point is not finite"))
                            continue
                        smp = job.samples() if callable(job.samples) else job.samples
                        ph, da, bk = smp.get("phsp"), smp.get("data"), smp.get("bkg")
                        n3 = (len(ph), len(da), 0 if bk is None else len(bk))
                        need = self.footprint(n3)
                    except Exception as exc:
                        ready.put(("failed", job, n3, "events: %s" % exc))
                        continue
                    if need > self.hbm_budget:
                        ready.put(("rejected", job, n3, "needs %d bytes of HBM, more than the service's budget of %d"
                                   % (need, self.hbm_budget)))
                        continue
                    with cond:
                        while mem["used"] + need > self.hbm_budget and not mem["stop"]:
                            cond.wait(0.05)
                        if mem["stop"]:
                            return
                        mem["used"] += need
                    ta = self.clock()
                    try:
                        h = _native.Handle(self.plan_text, ph, da, bk, smp.get("phsp_weights"),
                                           smp.get("data_weights"), smp.get("bkg_weights"), device=self.device,
                                           precision=self.precision)
                    except Exception as exc:
                        with cond:
                            mem["used"] -= need
                            cond.notify_all()
                        ready.put(("failed", job, n3, "handle: %s" % exc))
                        continue
                    st.admissions_seconds += self.clock() - ta
                    ready.put(("ok", job, h, need, sum(n3), self.clock() - t0))
            except BaseException as exc:          # surfaces in the main loop
                ready.put(exc)
                return
            ready.put(None)

        def refuse(job, n3, status, why):
            t = self.clock() - t0
            x0 = np.asarray(job.x0, dtype=np.float64)
            st.fits.append(FitRecord(job.job_id, Outcome(x0, float("nan"), 0, 0, False), int(sum(n3)), t, t,
                                     job.arrival, status, why))

        worker = threading.Thread(target=admit, daemon=True)
        worker.start()
        opt = None
        P = O = P_host = O_host = Hd = H_host = None
        active: Dict[int, dict] = {}
        free_slots = list(range(self.max_slots))[::-1]
        admitting = True
        try:
            while admitting or active:
                # fits whose handles are ready take the free slots
                while admitting and free_slots:
                    try:
                        item = ready.get(block=not active, timeout=None if not active else 0)
                    except _queue.Empty:
                        break
                    if item is None:
                        admitting = False
                        break
                    if isinstance(item, BaseException):
                        raise item
                    if item[0] != "ok":
                        refuse(item[1], item[2], item[0], item[3])
                        continue
                    _, job, h, need, nev, t_adm = item
                    if opt is None:
                        self.np = h.n_params
                        if self.method == "newton":
                            opt = BatchNewton(self.max_slots, self.np, edm_tol=self.edm_tol)
                            Hd = torch.zeros((self.max_slots, self.np, self.np), dtype=torch.float64, device=dev)
                            H_host = torch.zeros((self.max_slots, self.np, self.np), dtype=torch.float64).pin_memory()
                        else:
                            opt = BatchLBFGS(self.max_slots, self.np)
                        P = torch.zeros((self.max_slots, self.np), dtype=torch.float64, device=dev)
                        O = torch.zeros((self.max_slots, 1 + self.np), dtype=torch.float64, device=dev)
                        P_host = torch.zeros((self.max_slots, self.np), dtype=torch.float64).pin_memory()
                        O_host = torch.zeros((self.max_slots, 1 + self.np), dtype=torch.float64).pin_memory()
                    slot = free_slots.pop()
                    opt.start(slot, job.x0, job.lower, job.upper, job.max_evals)
                    active[slot] = dict(job=job, h=h, need=need, t_req=self.clock(), admitted=t_adm, n_events=nev,
                                        first=True)
                if not active:
                    continue
                # one round: every active fit's pending point in one batched evaluation
                tr = self.clock()
                slots = sorted(active)
                Ph = P_host.numpy()
                Ph[slots] = opt.xt[slots]
                failed: Dict[int, str] = {}
                with torch.cuda.stream(stream):
                    P.copy_(P_host, non_blocking=True)
                    try:
                        self.batch.nll_grad([active[s]["h"] for s in slots], [P[s].data_ptr() for s in slots],
                                            [O[s].data_ptr() for s in slots], stream.cuda_stream)
                    except RuntimeError:
                        # find the fits that fail: each evaluated alone (a fit's solo result
                        # is bit-identical to its batched one)
                        for s in slots:
                            try:
                                active[s]["h"].nll_grad(P[s].data_ptr(), O[s].data_ptr(), stream.cuda_stream)
                            except RuntimeError as exc:
                                failed[s] = "likelihood: %s" % exc
                    O_host.copy_(O, non_blocking=True)
                stream.synchronize()
                tdone = self.clock()
                st.round_seconds.append(tdone - tr)
                st.batch_sizes.append(len(slots))
                res = O_host.numpy()
                for s in slots:
                    # the reference reports a non-finite loss as 1e30 (src/Analysis.cu:1145-1158);
                    # at a fit's starting point that leaves nothing to minimise
                    if active[s]["first"] and s not in failed and not (abs(res[s, 0]) < 1e30):
                        failed[s] = "the likelihood is not finite at the starting point"
                    active[s]["first"] = False
                ok = [s for s in slots if s not in failed]
                done: Dict[int, Outcome] = {}
                if self.method == "newton":
                    done, need_h = opt.step(ok, res[ok, 0], res[ok, 1:]) if ok else ({}, [])
                    if need_h:
                        # exact Hessians where fits moved (P[s] holds the point just evaluated)
                        th = self.clock()
                        with torch.cuda.stream(stream):
                            for s in need_h:
                                try:
                                    active[s]["h"].hessian(P[s].data_ptr(), Hd[s].data_ptr(), stream.cuda_stream)
                                except RuntimeError as exc:
                                    failed[s] = "Hessian: %s" % exc
                            H_host.copy_(Hd, non_blocking=True)
                        stream.synchronize()
                        st.hessian_seconds.append(self.clock() - th)
                        need_h = [s for s in need_h if s not in failed]
                        if need_h:
                            done.update(opt.set_hessians(need_h, H_host.numpy()[need_h]))
                elif ok:
                    done = opt.step(ok, res[ok, 0], res[ok, 1:])
                tnext = self.clock()
                for s in ok:
                    a = active[s]
                    st.latencies.append(tdone - a["t_req"])
                    a["t_req"] = tnext

                def retire(s, outcome, status="ok", why=""):
                    a = active.pop(s)
                    st.fits.append(FitRecord(a["job"].job_id, outcome, a["n_events"], a["admitted"], tnext - t0,
                                             a["job"].arrival, status, why))
                    a["h"].close()
                    with cond:
                        mem["used"] -= a["need"]
                        cond.notify_all()
                    free_slots.append(s)

                for s, outcome in done.items():
                    if s not in failed:
                        retire(s, outcome)
                for s, why in failed.items():
                    x = np.asarray(active[s]["job"].x0, dtype=np.float64)
                    retire(s, Outcome(x, float("nan"), 0, 0, False), "failed", why)
        finally:
            with cond:
                mem["stop"] = True
                cond.notify_all()
            worker.join()
            for a in active.values():
                a["h"].close()
        st.wall = self.clock() - t0
        return st

    def close(self):
        self.batch.close()


# ---------------------------------------------------------------------------
# toy-MC ensemble (the serving workload)
# ---------------------------------------------------------------------------
def toy_ensemble(plan, phsp_pool: np.ndarray, data_pool: np.ndarray, n_jobs: int, seed: int,
                 n_min: int = 5_000, n_max: int = 500_000, phsp_factor: int = 5, rate: float = 20.0,
                 x_nominal: Optional[np.ndarray] = None, lower=None, upper=None, max_evals: int = 400) -> List[FitJob]:
    """Fits with data counts log-uniform in [n_min, n_max], phase space phsp_factor x data,
    Poisson arrivals at ``rate`` per second.  Each fit's events are a random window of the
    shared pools (taken at admission); couplings start from the nominal point (first
    coupling fixed to 1 + 0i) with a per-fit perturbation."""
    rng = np.random.default_rng(seed)
    n = plan.n_free
    x_nom = np.array(x_nominal, dtype=np.float64)
    v = x_nom[:n] + 1j * x_nom[n:2 * n]
    v = v / v[0]
    x_nom[:n], x_nom[n:2 * n] = v.real, v.imag
    lo = np.concatenate([np.full(2 * n, -1e4), np.array([t.lower for t in plan.theta])]) if lower is None else lower
    hi = np.concatenate([np.full(2 * n, 1e4), np.array([t.upper for t in plan.theta])]) if upper is None else upper
    lo, hi = lo.copy(), hi.copy()
    lo[0] = hi[0] = 1.0
    lo[n] = hi[n] = 0.0
    t = 0.0
    jobs = []
    for j in range(n_jobs):
        t += rng.exponential(1.0 / rate)
        nd = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        npp = min(phsp_factor * nd, len(phsp_pool))
        nd = min(nd, len(data_pool))
        d0 = int(rng.integers(0, len(data_pool) - nd + 1))
        p0 = int(rng.integers(0, len(phsp_pool) - npp + 1))
        x0 = x_nom.copy()
        x0[1:n] += 0.05 * rng.standard_normal(n - 1)
        x0[n + 1:2 * n] += 0.05 * rng.standard_normal(n - 1)
        x0[2 * n:] *= 1.0 + 0.01 * rng.standard_normal(len(x0) - 2 * n)
        x0 = np.clip(x0, lo, hi)

        def take(d0=d0, nd=nd, p0=p0, npp=npp):
            return {"data": data_pool[d0:d0 + nd], "phsp": phsp_pool[p0:p0 + npp]}
        jobs.append(FitJob(j, take, x0, lo, hi, arrival=t, max_evals=max_evals))
    return jobs
