"""The two-family four-body cascade on the GPU (J/psi -> K+ K- pi+ pi-: K1(1270), K1(1400),
K1(1650) -> K*(892) pi and K rho(770), K*0 K*0bar, K*0 K*2(1430)0bar and K*0 K*0(1430)0bar over
three K* states, phi or f2'(1525) with pi+ pi- resonances; 149 amplitudes, 24 free resonance
parameters, a 322 x 322 Hessian).

* Reduced size (8k phase space, 2k data): NLL, gradient and Hessian against the oracle,
  FP64 form 1e-9 / 1e-7 / 1e-6 relative; the mixed form's Hessian (FP32 Jacobian on the
  FP32 matrix cores, FP64 sums across tiles) within 1e-5.
* Full size (10M phase space, 2M phase-space distributed data): the Hessian is symmetric to 1e-9 of its
  largest entry, and its rows match central differences of the GPU gradient (steps of
  1e-5 relative) to 1e-4 of each row's largest entry, for couplings and resonance
  parameters of both chain families.
"""
import numpy as np
import pytest
import torch

from helpers import perturbed, plan_of, rel, samples
from ctpwa_amd import workloads as W

pytestmark = pytest.mark.gpu

SPEC = W.WORKLOADS["jpsi_kkpipi_wide"]
MASSES = SPEC["masses"]


def _run(h, p, hess=False):
    t = torch.tensor(p, dtype=torch.float64, device="cuda")
    out = torch.empty(1 + h.n_params, dtype=torch.float64, device="cuda")
    h.nll_grad(t.data_ptr(), out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    H = None
    if hess:
        Ht = torch.empty((h.n_params, h.n_params), dtype=torch.float64, device="cuda")
        h.hessian(t.data_ptr(), Ht.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    if hess:
        H = Ht.cpu().numpy()
    return o[0], o[1:], H


def test_wide_counts():
    plan = plan_of("jpsi_kkpipi_wide")
    assert (plan.n_amp, plan.n_free, plan.n_theta, plan.npol) == (149, 149, 24, 2)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_wide_against_oracle(oracle_lib, precision):
    from ctpwa_amd import _native
    plan = plan_of("jpsi_kkpipi_wide")
    ph, da, _ = samples(8000, 2000, 0, 81, masses=MASSES)
    print("case jpsi_kkpipi_wide 8000 2000", precision, flush=True)
    h = _native.Handle(plan.to_text(), ph, da, None, precision=precision)
    O = oracle_lib.Oracle(plan.to_text(), 4, ph, da)
    p = perturbed(W.nominal_params(plan, 4), plan, 5)
    nll, g, H = _run(h, p, hess=True)
    nll_o, g_o, H_o = O.hessian(p)
    tol = (1e-9, 1e-7, 1e-6) if precision == "fp64" else (1e-6, 1e-5, 1e-5)
    print("wide", precision, nll, nll_o, rel(g, g_o), rel(H, H_o), flush=True)
    assert abs(nll - nll_o) <= tol[0] * abs(nll_o)
    assert rel(g, g_o) <= tol[1]
    assert rel(H, H_o) <= tol[2]
    h.close()


def test_wide_full_size_hessian():
    from ctpwa_amd import _native
    plan = plan_of("jpsi_kkpipi_wide")
    print("case jpsi_kkpipi_wide full size: generating", flush=True)
    # derivative checks: phase-space distributed data serve as well as the model's
    ph = W.phase_space(SPEC["n_phsp"], seed=1004, masses=MASSES, workers=8)
    da = W.phase_space(SPEC["n_data"], seed=3004, masses=MASSES, workers=8)
    print("case jpsi_kkpipi_wide full size: handle", flush=True)
    h = _native.Handle(plan.to_text(), ph, da, None)
    del ph, da
    p = perturbed(W.nominal_params(plan, 4), plan, 23)
    _, g, H = _run(h, p, hess=True)
    print("full-size Hessian done", flush=True)
    assert np.all(np.isfinite(H))
    assert np.max(np.abs(H - H.T)) <= 1e-9 * np.max(np.abs(H))
    n = plan.n_free
    # couplings of the K1 -> K* pi, K1 -> K rho and K* K*bar families; K1(1400), K*(1410),
    # rho(770) and f0(980) parameters
    names = [t.name for t in plan.theta]
    picks = [1, 40, 120, n + 7, n + 100]
    for want in ("K1_1650", "Kst1410", "rho1450", "f0_980"):
        picks += [2 * n + i for i, nm in enumerate(names) if nm.startswith(want)][:1]
    for j in picks:
        step = 1e-5 * max(abs(p[j]), 1e-2)
        pp, pm = p.copy(), p.copy()
        pp[j] += step
        pm[j] -= step
        _, gp, _ = _run(h, pp)
        _, gm, _ = _run(h, pm)
        fdh = (gp - gm) / (2 * step)
        err = np.max(np.abs(fdh - H[j])) / np.max(np.abs(H[j]))
        print("fd row", j, err, flush=True)
        assert err <= 1e-4
    h.close()
