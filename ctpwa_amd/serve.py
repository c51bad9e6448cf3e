"""Resident fit service with continuous batching of NLL+gradient requests.

Many fits of one model (multi-start minimisations, toy-MC ensembles with uneven event
counts) share the GPU.  Every fit is a minimiser written as a generator: it yields the
next parameter point and receives (NLL, gradient), so the service drives all active fits
in lock step without threads:

    round:  gather the pending point of every active fit
            -> one batched evaluation (ctpwa_batch_nll_grad: batched prep, one grouped
               per-event launch whose workgroups are shared out among the fits, batched
               reduction)
            -> hand each fit its result; finished fits retire and free their slot and HBM
    between rounds: admit arrived fits while a slot is free and their device footprint
    (ctpwa_footprint_bytes) fits the HBM budget.

Each fit's events stay resident on the device for its lifetime.  A fit's results are
bit-identical whether it runs alone or in a batch (the batched evaluation reproduces the
single one exactly), so a fit's trajectory does not depend on what else is running.

The reference has no service; its fit driver (example/fit.py:84-260) runs one fit per
process with torch's L-BFGS.  The minimiser here is a bounded L-BFGS (two-loop recursion,
projected gradient, Armijo backtracking) kept on the host in float64 numpy.
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
    job_id: int
    outcome: Outcome
    n_events: int
    admitted: float
    finished: float


@dataclass
class ServiceStats:
    fits: List[FitRecord] = field(default_factory=list)
    latencies: List[float] = field(default_factory=list)     # per evaluation, seconds
    batch_sizes: List[int] = field(default_factory=list)
    round_seconds: List[float] = field(default_factory=list)
    wall: float = 0.0
    admissions_seconds: float = 0.0

    def summary(self) -> Dict[str, float]:
        lat = np.array(self.latencies) if self.latencies else np.zeros(1)
        return {
            "fits": len(self.fits),
            "evaluations": int(len(self.latencies)),
            "wall_s": self.wall,
            "fits_per_hour": len(self.fits) / self.wall * 3600.0 if self.wall > 0 else 0.0,
            "evals_per_s": len(self.latencies) / self.wall if self.wall > 0 else 0.0,
            "latency_p50_ms": float(np.percentile(lat, 50) * 1e3),
            "latency_p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_batch": float(np.mean(self.batch_sizes)) if self.batch_sizes else 0.0,
            "rounds": len(self.batch_sizes),
            "admission_s": self.admissions_seconds,
            "converged": int(sum(r.outcome.converged for r in self.fits)),
        }


class FitService:
    """Runs fits of one model (``plan_text``) with continuous batching on one GPU."""

    def __init__(self, plan_text: str, precision="fp64", max_slots: int = 64, hbm_budget: Optional[int] = None,
                 device: int = 0, clock: Callable[[], float] = time.perf_counter):
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
        self.stream = torch.cuda.current_stream(device)

    def footprint(self, n3) -> int:
        return _native.footprint_bytes(self.plan_text, n3[0], n3[1], n3[2], self.precision)

    def run(self, jobs: List[FitJob]) -> ServiceStats:
        st = ServiceStats()
        queue = sorted(jobs, key=lambda j: j.arrival)
        qi = 0
        active: Dict[int, dict] = {}           # slot -> state
        free_slots = list(range(self.max_slots))[::-1]
        mem_used = 0
        dev = torch.device("cuda", self.device)
        P = O = P_host = O_host = None
        t0 = self.clock()
        while qi < len(queue) or active:
            now = self.clock() - t0
            # admission: arrived fits while a slot and memory are free
            while qi < len(queue) and queue[qi].arrival <= now and free_slots:
                job = queue[qi]
                smp = job.samples() if callable(job.samples) else job.samples
                ph, da, bk = smp.get("phsp"), smp.get("data"), smp.get("bkg")
                n3 = (len(ph), len(da), 0 if bk is None else len(bk))
                need = self.footprint(n3)
                if need > self.hbm_budget:
                    raise RuntimeError("fit %d needs %d bytes, more than the HBM budget" % (job.job_id, need))
                if mem_used + need > self.hbm_budget:
                    break
                ta = self.clock()
                h = _native.Handle(self.plan_text, ph, da, bk, smp.get("phsp_weights"), smp.get("data_weights"),
                                   smp.get("bkg_weights"), device=self.device, precision=self.precision)
                st.admissions_seconds += self.clock() - ta
                if self.np is None:
                    self.np = h.n_params
                    P = torch.zeros((self.max_slots, self.np), dtype=torch.float64, device=dev)
                    O = torch.zeros((self.max_slots, 1 + self.np), dtype=torch.float64, device=dev)
                    P_host = torch.zeros((self.max_slots, self.np), dtype=torch.float64).pin_memory()
                    O_host = torch.zeros((self.max_slots, 1 + self.np), dtype=torch.float64).pin_memory()
                gen = lbfgs(job.x0, job.lower, job.upper, max_evals=job.max_evals)
                x = next(gen)
                slot = free_slots.pop()
                mem_used += need
                active[slot] = dict(job=job, h=h, gen=gen, x=x, t_req=self.clock(), need=need,
                                    admitted=self.clock() - t0, n_events=sum(n3))
                qi += 1
                now = self.clock() - t0
            if not active:
                if qi < len(queue):
                    wait = queue[qi].arrival - (self.clock() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.01))
                continue
            # one round: every active fit's pending point in one batched evaluation
            tr = self.clock()
            slots = sorted(active)
            for s in slots:
                P_host[s].numpy()[:] = active[s]["x"]
            P.copy_(P_host, non_blocking=True)
            self.batch.nll_grad([active[s]["h"] for s in slots], [P[s].data_ptr() for s in slots],
                                [O[s].data_ptr() for s in slots], self.stream.cuda_stream)
            O_host.copy_(O, non_blocking=True)
            self.stream.synchronize()
            tdone = self.clock()
            st.round_seconds.append(tdone - tr)
            st.batch_sizes.append(len(slots))
            res = O_host.numpy()
            for s in slots:
                a = active[s]
                st.latencies.append(tdone - a["t_req"])
                try:
                    a["x"] = a["gen"].send((float(res[s, 0]), res[s, 1:].copy()))
                    a["t_req"] = self.clock()
                except StopIteration as fin:
                    st.fits.append(FitRecord(a["job"].job_id, fin.value, a["n_events"], a["admitted"],
                                             self.clock() - t0))
                    a["h"].close()
                    mem_used -= a["need"]
                    free_slots.append(s)
                    del active[s]
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
