"""trans and var_equal through the likelihood (oracle, CPU).

A constrained plan is a reparametrisation of the unconstrained one: with trans [A, B]: r
the couplings of B's amplitudes are r times those of A's, with var_equal the members'
parameters equal the owner's.  So NLL_constrained(p) = NLL_free(E p) for the expansion E,
and grad_constrained = E^T grad_free, Hessian likewise.  Checked on the oracle for both
constraints together (the GPU is compared with the oracle on the same plans in the GPU
suite).
"""
import numpy as np

from helpers import plan_of, samples, perturbed
from ctpwa_amd import workloads as W

TEXT = (W.JPSI_3PI + "Constraints:\n  trans:\n    - [R_m0, R_p0]: -1\n"
        "  var_equal:\n    - [rho1450_width, rho1700_width]\n")


def _expansion(pc, pf):
    """Matrix E with x_free = E x_constrained (grouped layouts)."""
    nc, nf = pc.n_free, pf.n_free
    E = np.zeros((2 * nf + pf.n_theta, 2 * nc + pc.n_theta))
    # couplings: every free-plan amplitude has its own coupling (no split, one amp per chain)
    assert pf.amp_free == list(range(pf.n_amp))
    for a in range(pf.n_amp):
        E[a, pc.amp_free[a]] = pc.amp_ratio[a]
        E[nf + a, nc + pc.amp_free[a]] = pc.amp_ratio[a]
    cn = [t.name for t in pc.theta]
    for j, t in enumerate(pf.theta):
        src = "rho1450_width" if t.name == "rho1700_width" else t.name
        E[2 * nf + j, 2 * nc + cn.index(src)] = 1.0
    return E


def test_constrained_equals_expanded(oracle_lib):
    pc = plan_of(TEXT)
    pf = plan_of("jpsi_3pi")
    assert (pc.n_free, pc.n_theta) == (7, 9)
    ph, da, bk = samples(6000, 1500, 400, 5)
    wb = np.full(bk.shape[0], 0.8)
    Oc = oracle_lib.Oracle(pc.to_text(), 3, ph, da, bk, w_bkg=wb)
    Of = oracle_lib.Oracle(pf.to_text(), 3, ph, da, bk, w_bkg=wb)
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(-1, 1, 2 * pc.n_free), [t.init for t in pc.theta]])
    x[2 * pc.n_free:] *= 1.0 + 0.01 * rng.standard_normal(pc.n_theta)
    E = _expansion(pc, pf)
    nll_c, g_c, H_c = Oc.hessian(x)
    nll_f, g_f, H_f = Of.hessian(E @ x)
    assert abs(nll_c - nll_f) <= 1e-12 * abs(nll_f)
    assert np.allclose(g_c, E.T @ g_f, rtol=1e-10, atol=1e-9 * np.max(np.abs(g_f)))
    assert np.allclose(H_c, E.T @ H_f @ E, rtol=1e-9, atol=1e-9 * np.max(np.abs(H_f)))
