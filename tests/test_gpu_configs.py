"""The benchmark configurations at their stated sizes against the CPU FP64 oracle.

* J/psi -> rho(770) pi0 (one chain, BWR mass and width free) exactly as specified for the
  CPU-evaluable case: 50k GENBOD phase-space events from seed 0 and 10k data events by
  accept-reject on the model intensity (both J/psi spin projections, summed
  incoherently), no background.  The accept-reject intensity is the oracle's, so the
  events do not depend on the code under test.  NLL, gradient and Hessian in the FP64
  form (bars 1e-9, 1e-7, 1e-6 relative) and in the mixed form (FP32 amplitude tensors:
  NLL 1e-9 relative, gradient and Hessian 1e-5 relative, the storage rounding bound).
* J/psi -> pi+ pi- pi0 with 3 chains x (rho(770), rho(1450), rho(1700)) + f0(980) +
  f2(1270) at full size: 500k data by accept-reject on the model, 2.5M phase space,
  50k background at weight 0.8.  NLL and gradient at two points in both forms and the
  FP64 Hessian at one point, against the oracle run on all host cores (a few seconds per
  evaluation).  Bars: FP64 as above; mixed |Delta NLL| <= 1e-4 (absolute) and gradient
  1e-5 relative.
"""
import os

import numpy as np
import pytest
import torch

from helpers import perturbed, plan_of, rel
from ctpwa_amd import workloads as W

pytestmark = pytest.mark.gpu


def _oracle_intensity(oracle_lib):
    def intensity(plan_text, events, params):
        O = oracle_lib.Oracle(plan_text, events.shape[1], events, events[:1], None, threads=_threads())
        return O.intensity(params, 0)
    return intensity


def _threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _eval(h, p, hess=False):
    t = torch.tensor(p, dtype=torch.float64, device="cuda")
    out = torch.empty(1 + h.n_params, dtype=torch.float64, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    h.nll_grad(t.data_ptr(), out.data_ptr(), s)
    H = None
    if hess:
        Ht = torch.empty((h.n_params, h.n_params), dtype=torch.float64, device="cuda")
        h.hessian(t.data_ptr(), Ht.data_ptr(), s)
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    if hess:
        H = Ht.cpu().numpy()
    return o[0], o[1:], H


def test_single_chain_config_as_specified(oracle_lib):
    from ctpwa_amd import _native
    name = "jpsi_rho"
    spec = W.WORKLOADS[name]
    assert (spec["n_phsp"], spec["n_data"], spec["n_bkg"], spec["seed"]) == (50_000, 10_000, 0, 0)
    s = W.generate_samples(name, _oracle_intensity(oracle_lib))
    ph, da = s["phsp"], s["data"]
    assert ph.shape == (50_000, 3, 4) and da.shape == (10_000, 3, 4) and s["bkg"] is None
    plan = plan_of(name)
    print("case", name, "as specified", ph.shape[0], da.shape[0], flush=True)
    O = oracle_lib.Oracle(plan.to_text(), 3, ph, da, None, threads=_threads())
    h64 = _native.Handle(plan.to_text(), ph, da, None)
    hmx = _native.Handle(plan.to_text(), ph, da, None, precision="mixed")
    base = W.nominal_params(plan, None)
    for k, p in enumerate([base, perturbed(base, plan, 31, 0.2)]):
        nll_o, g_o, H_o = O.hessian(p)
        nll_g, g_g, H_g = _eval(h64, p, hess=True)
        nll_m, g_m, H_m = _eval(hmx, p, hess=True)
        print(k, "fp64", nll_o, nll_g, rel(nll_g, nll_o), rel(g_g, g_o), rel(H_g, H_o), flush=True)
        print(k, "mixed", nll_m - nll_o, rel(g_m, g_o), rel(H_m, H_o), flush=True)
        assert abs(nll_g - nll_o) <= 1e-9 * abs(nll_o)
        assert rel(g_g, g_o) <= 1e-7
        assert rel(H_g, H_o) <= 1e-6
        assert abs(nll_m - nll_o) <= 1e-9 * abs(nll_o)
        assert rel(g_m, g_o) <= 1e-5
        assert rel(H_m, H_o) <= 1e-5


@pytest.fixture(scope="module")
def full_c1():
    """The three-pion workload at full size (data and background by accept-reject on the
    model intensity on the device, as the benchmark draws them)."""
    spec = W.WORKLOADS["jpsi_3pi"]
    s = W.generate_samples("jpsi_3pi", W.device_intensity, n_data=spec["n_data"] + spec["n_bkg"], n_bkg=0)
    return s["phsp"], s["data"][:spec["n_data"]], s["data"][spec["n_data"]:]


def test_three_pion_full_size_vs_oracle(full_c1, oracle_lib):
    from ctpwa_amd import _native
    ph, da, bk = full_c1
    assert (ph.shape[0], da.shape[0], bk.shape[0]) == (2_500_000, 500_000, 50_000)
    plan = plan_of("jpsi_3pi")
    wb = np.full(bk.shape[0], 0.8)
    print("case jpsi_3pi full size", ph.shape[0], da.shape[0], bk.shape[0], "oracle threads", _threads(), flush=True)
    O = oracle_lib.Oracle(plan.to_text(), 3, ph, da, bk, w_bkg=wb, threads=_threads())
    h64 = _native.Handle(plan.to_text(), ph, da, bk, None, None, wb)
    hmx = _native.Handle(plan.to_text(), ph, da, bk, None, None, wb, precision="mixed")
    base = perturbed(W.nominal_params(plan, 1), plan, 41)
    for k, p in enumerate([base, perturbed(base, plan, 42, 0.2)]):
        hess = k == 0
        if hess:
            nll_o, g_o, H_o = O.hessian(p)
        else:
            nll_o, g_o = O.nll_grad(p)
        nll_g, g_g, H_g = _eval(h64, p, hess=hess)
        nll_m, g_m, _ = _eval(hmx, p)
        print(k, "fp64", nll_o, nll_g, rel(nll_g, nll_o), rel(g_g, g_o),
              rel(H_g, H_o) if hess else "-", flush=True)
        print(k, "mixed", nll_m - nll_o, rel(g_m, g_o), flush=True)
        assert abs(nll_g - nll_o) <= 1e-9 * abs(nll_o)
        assert rel(g_g, g_o) <= 1e-7
        if hess:
            assert rel(H_g, H_o) <= 1e-6
        assert abs(nll_m - nll_o) <= 1e-4
        assert rel(g_m, g_o) <= 1e-5
