"""No free resonance parameter: the phase-space term of every evaluation comes from the
amplitude Gram stored at creation (c^H G c for the integral, G c for its gradient sums;
the reference's stored phase-space matrix and quadratic form, src/Analysis.cu:808-826,
src/ComputeNLL.cu:575-597).

Checked against the CPU oracle (FP64 bars of the other parity tests: NLL 1e-9, gradient
1e-7 relative), against the per-event phase-space scan of the same inputs
(CTPWA_FIXED_GRAM=0) at the full jpsi_3pi size, and through the batched and no-alloc
entry points, whose results must equal the single evaluation's bit for bit."""
import os

import numpy as np
import pytest
import torch

from helpers import perturbed, plan_of, rel, samples
from ctpwa_amd import workloads as W

pytestmark = pytest.mark.gpu

# jpsi_3pi with every resonance mass and width fixed at its configured value
FIXED_3PI = W.fixed_theta_config(W.JPSI_3PI)


def _plan():
    plan = plan_of(FIXED_3PI)
    assert plan.n_theta == 0 and plan.n_free > 0
    return plan


def _handle(plan, ph, da, bk, wb, precision, scan):
    from ctpwa_amd import _native
    old = os.environ.get("CTPWA_FIXED_GRAM")
    if scan:
        os.environ["CTPWA_FIXED_GRAM"] = "0"
    try:
        return _native.Handle(plan.to_text(), ph, da, bk, None, None, wb, precision=precision)
    finally:
        if old is None:
            os.environ.pop("CTPWA_FIXED_GRAM", None)
        else:
            os.environ["CTPWA_FIXED_GRAM"] = old


def _eval(h, p, noalloc=False):
    t = torch.tensor(p, dtype=torch.float64, device="cuda")
    out = torch.empty(1 + h.n_params, dtype=torch.float64, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    (h.nll_grad_noalloc if noalloc else h.nll_grad)(t.data_ptr(), out.data_ptr(), s)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _points(plan):
    base = W.nominal_params(plan, 1)
    return [base, perturbed(base, plan, 5), perturbed(base, plan, 6, 0.3)]


def test_layout_skips_phase_space():
    plan = _plan()
    ph, da, bk = samples(20000, 4000, 1000, 41)
    wb = np.full(bk.shape[0], 0.8)
    hg = _handle(plan, ph, da, bk, wb, "fp64", scan=False)
    hs = _handle(plan, ph, da, bk, wb, "fp64", scan=True)
    lg, ls = hg.layout(), hs.layout()
    print("layouts", lg, ls, flush=True)
    per_event = ls["bytes_per_eval"] // sum(ls["n_events"])
    assert lg["bytes_per_eval"] == (4000 + 1000) * per_event
    assert lg["n_workgroups"] < ls["n_workgroups"]
    # a plan with a free resonance parameter keeps the scan
    hf = _handle(plan_of("jpsi_3pi"), ph, da, bk, wb, "fp64", scan=False)
    lf = hf.layout()
    assert lf["bytes_per_eval"] >= sum(lf["n_events"]) * 16


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fixed_theta_against_oracle_and_scan(precision, oracle_lib):
    plan = _plan()
    ph, da, bk = samples(30000, 6000, 1500, 43)
    wb = np.full(bk.shape[0], 0.8)
    print("case fixed-theta jpsi_3pi 30000/6000/1500", precision, flush=True)
    hg = _handle(plan, ph, da, bk, wb, precision, scan=False)
    hs = _handle(plan, ph, da, bk, wb, precision, scan=True)
    O = oracle_lib.Oracle(plan.to_text(), 3, ph, da, bk, w_bkg=wb)
    for k, p in enumerate(_points(plan)):
        g = _eval(hg, p)
        s = _eval(hs, p)
        nll_o, g_o = O.nll_grad(p)
        print(k, nll_o, g[0], s[0], rel(g[0], nll_o), rel(g[1:], g_o), rel(g, s), flush=True)
        # the stored Gram against the scan of the same stored inputs
        assert abs(g[0] - s[0]) <= 1e-11 * abs(s[0])
        assert rel(g[1:], s[1:]) <= 1e-9
        if precision == "fp64":
            assert abs(g[0] - nll_o) <= 1e-9 * abs(nll_o)
            assert rel(g[1:], g_o) <= 1e-7
        else:
            # FP32 storage rounding (2^-24 per stored value; the scan shows the same
            # deviation, the Gram and scan agree to 1e-11 above)
            assert abs(g[0] - nll_o) <= 1e-8 * abs(nll_o)
            assert rel(g[1:], g_o) <= 1e-5
        # the no-alloc entry point on the same handle: the same kernels, the same bits
        assert np.array_equal(_eval(hg, p, noalloc=True), g)


def test_fixed_theta_full_size_batch_and_rate():
    """The full jpsi_3pi event counts (2.5M phase space, 500k data, 50k background at
    weight 0.8), mixed form: the stored Gram agrees with the scan, a batch holding a Gram
    handle and a scanning handle of one model returns each one's single evaluation bit for
    bit, and the fused kernel reads only the data and background events."""
    from ctpwa_amd import _native
    cfg = W.WORKLOADS["jpsi_3pi"]
    plan = _plan()
    ph, da, bk = samples(cfg["n_phsp"], cfg["n_data"], cfg["n_bkg"], 47)
    wb = np.full(bk.shape[0], 0.8)
    print("case fixed-theta jpsi_3pi full size mixed", flush=True)
    hg = _handle(plan, ph, da, bk, wb, "mixed", scan=False)
    hs = _handle(plan, ph, da, bk, wb, "mixed", scan=True)
    pts = _points(plan)
    solo_g = [_eval(hg, p) for p in pts]
    solo_s = [_eval(hs, p) for p in pts]
    for g, s in zip(solo_g, solo_s):
        print("full", g[0], s[0], rel(g, s), flush=True)
        assert abs(g[0] - s[0]) <= 1e-10 * abs(s[0])
        assert rel(g[1:], s[1:]) <= 1e-8
    # a handle holds one table and one set of partials: each job of a batch has its own
    hg2 = _handle(plan, ph, da, bk, wb, "mixed", scan=False)
    hs2 = _handle(plan, ph, da, bk, wb, "mixed", scan=True)
    B = _native.Batch(4)
    ps = [torch.tensor(p, dtype=torch.float64, device="cuda") for p in pts[:2]]
    outs = [torch.empty(1 + hg.n_params, dtype=torch.float64, device="cuda") for _ in range(4)]
    with pytest.raises(_native.NativeError):
        B.nll_grad([hg, hg], [ps[0].data_ptr()] * 2, [o.data_ptr() for o in outs[:2]],
                   torch.cuda.current_stream().cuda_stream)
    assert "twice" in _native.last_error()
    hs_list = [hg, hs, hs2, hg2]
    pp = [ps[0], ps[0], ps[1], ps[1]]
    for rep in range(2):
        B.nll_grad(hs_list, [x.data_ptr() for x in pp], [o.data_ptr() for o in outs],
                   torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        got = [o.cpu().numpy() for o in outs]
        assert np.array_equal(got[0], solo_g[0]), rep
        assert np.array_equal(got[1], solo_s[0]), rep
        assert np.array_equal(got[2], solo_s[1]), rep
        assert np.array_equal(got[3], solo_g[1]), rep
    B.close()
    t = torch.tensor(pts[1], dtype=torch.float64, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    ms_g = hg.time_eval_kernel(t.data_ptr(), 50, s)
    ms_s = hs.time_eval_kernel(t.data_ptr(), 50, s)
    print("fused kernel ms: stored Gram %.4f, scan %.4f" % (ms_g, ms_s), flush=True)
    assert hg.layout()["bytes_per_eval"] * 5 < hs.layout()["bytes_per_eval"]
