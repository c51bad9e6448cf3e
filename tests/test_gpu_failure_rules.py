"""The reference's numeric failure rules, on the GPU against the oracle.

* I = 0 (src/ComputeNLL.cu:40-66): an event of zero intensity adds -w ln(1e-10) to the
  data log-sum and nothing to the gradient.  With every coupling zero all events have
  I = 0 and the phase-space mean pf = 0 is floored at 1e-30, so the loss has a closed
  form, -W_d ln 1e-10 + W_b ln 1e-10 + (W_d - W_b) ln 1e-30, with a zero gradient.
* w clamp (same lines): the per-event gradient weight w_e S_p / I is clamped to +-1e10
  componentwise.  Events of weight 1e16 hit it; the returned gradient then differs from
  the derivative of the loss (central differences), and the GPU matches the oracle.
* NaN / Inf loss (src/Analysis.cu:1145-1158): reported as 1e30 with a zero gradient.
"""
import math

import numpy as np
import pytest
import torch

from helpers import plan_of, rel, samples
from test_gpu_api import _write_cfg
from ctpwa_amd import workloads as W

pytestmark = pytest.mark.gpu


def _gpu(ana, p):
    t = torch.tensor(p, dtype=torch.float64, device="cuda", requires_grad=True)
    nll = ana.getNLL(t)
    (g,) = torch.autograd.grad(nll, t)
    return float(nll.item()), g.cpu().numpy()


def test_zero_intensity_rule(tmp_path, oracle_lib):
    import ctpwa_amd as ctpwa
    ph, da, bk = samples(6000, 1500, 400, 61)
    cfg = _write_cfg(tmp_path, "jpsi_3pi", ph, da, bk)
    print("case failure rules: I = 0", flush=True)
    plan = plan_of("jpsi_3pi")
    p = W.nominal_params(plan, 1)
    p[: 2 * plan.n_free] = 0.0
    closed = (-da.shape[0] * math.log(1e-10) + 0.8 * bk.shape[0] * math.log(1e-10)
              + (da.shape[0] - 0.8 * bk.shape[0]) * math.log(1e-30))
    for prec in ("fp64", "mixed"):
        ana = ctpwa.analysis(cfg, precision=prec)
        nll, g = _gpu(ana, p)
        assert abs(nll - closed) <= 1e-12 * abs(closed), (prec, nll, closed)
        assert np.all(g == 0.0), (prec, g)
    O = oracle_lib.Oracle(plan.to_text(), 3, ph, da, bk, w_bkg=np.full(bk.shape[0], 0.8))
    nll_o, g_o = O.nll_grad(p)
    assert abs(nll_o - closed) <= 1e-12 * abs(closed) and np.all(g_o == 0.0)


def test_gradient_weight_clamp(tmp_path, oracle_lib):
    import ctpwa_amd as ctpwa
    ph, da, _ = samples(6000, 1500, 0, 62)
    w = np.ones(da.shape[0])
    w[[3, 50, 400, 901, 1200]] = 1e16
    cfg = _write_cfg(tmp_path, "jpsi_3pi", ph, da, None)
    print("case failure rules: w clamp", flush=True)
    ana = ctpwa.analysis(cfg, samples={"phsp": ph, "data": da, "data_weights": w})
    plan = ana.plan
    O = oracle_lib.Oracle(plan.to_text(), 3, ph, da, w_data=w)
    p = W.nominal_params(plan, 1)
    nll, g = _gpu(ana, p)
    nll_o, g_o = O.nll_grad(p)
    assert abs(nll - nll_o) <= 1e-9 * abs(nll_o)
    assert rel(g, g_o) <= 1e-7
    # the clamp is active: the coupling gradient is not the derivative of the loss
    n = plan.n_free
    k = 1
    h = 1e-6 * max(1.0, abs(p[k]))
    pp, pm = p.copy(), p.copy()
    pp[k] += h
    pm[k] -= h
    fd = (O.nll_grad(pp)[0] - O.nll_grad(pm)[0]) / (2 * h)
    assert abs(fd - g_o[k]) > 1e-2 * abs(fd), (fd, g_o[k])
    # and without the heavy events the same comparison agrees (the clamp is what differs)
    O1 = oracle_lib.Oracle(plan.to_text(), 3, ph, da)
    fd1 = (O1.nll_grad(pp)[0] - O1.nll_grad(pm)[0]) / (2 * h)
    assert abs(fd1 - O1.nll_grad(p)[1][k]) <= 1e-5 * abs(fd1)
    assert n > 1


@pytest.mark.parametrize("what", ["nan_theta", "inf_coupling"])
def test_non_finite_loss_rule(tmp_path, oracle_lib, what):
    import ctpwa_amd as ctpwa
    ph, da, bk = samples(4000, 1000, 300, 63)
    cfg = _write_cfg(tmp_path, "jpsi_3pi", ph, da, bk)
    print("case failure rules:", what, flush=True)
    ana = ctpwa.analysis(cfg)
    plan = ana.plan
    p = W.nominal_params(plan, 1)
    if what == "nan_theta":
        p[2 * plan.n_free + 1] = np.nan
    else:
        p[2] = np.inf
    nll, g = _gpu(ana, p)
    assert nll == 1e30 and np.all(g == 0.0), (nll, g)
    O = oracle_lib.Oracle(plan.to_text(), 3, ph, da, bk, w_bkg=np.full(bk.shape[0], 0.8))
    nll_o, g_o = O.nll_grad(p)
    assert nll_o == 1e30 and np.all(g_o == 0.0)
