"""Shared workload fixtures for the tests (seeded synthetic samples)."""
import numpy as np

from ctpwa_amd import workloads as W
from ctpwa_amd.config import ConfigParser
from ctpwa_amd.model import build_plan


def plan_of(name_or_text):
    text = W.WORKLOADS[name_or_text]["config"] if name_or_text in W.WORKLOADS else name_or_text
    return build_plan(ConfigParser("", text=text))


def samples(n_phsp, n_data, n_bkg, seed, masses=None):
    ph = W.phase_space(n_phsp, seed=seed, masses=masses)
    da = W.phase_space(n_data, seed=seed + 1, masses=masses)
    bk = W.phase_space(n_bkg, seed=seed + 2, masses=masses) if n_bkg else None
    return ph, da, bk


def perturbed(params, plan, seed, scale=0.05):
    rng = np.random.default_rng(seed)
    p = params.copy()
    n = plan.n_free
    p[: 2 * n] += scale * rng.standard_normal(2 * n)
    t = p[2 * n:]
    t *= 1.0 + 0.02 * rng.standard_normal(t.shape)
    return p


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
