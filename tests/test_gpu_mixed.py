"""Mixed-precision form (FP32 SL tensors and invariants, FP64 squared masses, FP64
arithmetic and sums) against the oracle and against the FP64 form.

Gates: |Delta NLL| <= 1e-4 at the benchmark size (500k data, 2.5M phase space, 50k
background), fitted parameters within 0.01 sigma of the FP64 fit (sigma from the FP64
Hessian).  At small sizes the storage rounding (2^-24 relative per stored value) bounds
the deviation: NLL 1e-9 relative, gradient and Hessian 1e-5 relative, amplitudes 1e-6 of
the largest amplitude.
"""
import numpy as np
import pytest
import torch

from helpers import perturbed, plan_of, rel, samples
from ctpwa_amd import workloads as W

pytestmark = pytest.mark.gpu


def _ng(h, p, hess=False):
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


@pytest.mark.parametrize("name", ["jpsi_3pi", "jpsi_kkpipi"])
def test_mixed_small_vs_oracle(name, oracle_lib):
    from ctpwa_amd import _native
    spec = W.WORKLOADS[name]
    plan = plan_of(name)
    nf = len(plan.final_order)
    nb = 1500 if spec["n_bkg"] else 0
    ph, da, bk = samples(30000, 6000, nb, 81, masses=spec.get("masses"))
    wb = None if bk is None else np.full(bk.shape[0], 0.8)
    print("case mixed", name, flush=True)
    h = _native.Handle(plan.to_text(), ph, da, bk, None, None, wb, precision="mixed")
    O = oracle_lib.Oracle(plan.to_text(), nf, ph, da, bk, w_bkg=wb)
    p = perturbed(W.nominal_params(plan, spec["seed"]), plan, 3)
    nll_o, g_o, H_o = O.hessian(p)
    nll_m, g_m, H_m = _ng(h, p, hess=True)
    print(name, nll_o, nll_m, nll_m - nll_o, rel(g_m, g_o), rel(H_m, H_o), flush=True)
    assert abs(nll_m - nll_o) <= 1e-9 * abs(nll_o)
    assert rel(g_m, g_o) <= 1e-5
    assert rel(H_m, H_o) <= 1e-5
    t = torch.tensor(p, dtype=torch.float64, device="cuda")
    n = h.n_events[1]
    out = torch.empty((n, plan.npol, plan.n_amp, 2), dtype=torch.float64, device="cuda")
    h.amplitudes(t.data_ptr(), 1, out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    A = out.cpu().numpy()
    A = A[..., 0] + 1j * A[..., 1]
    Ao = O.amplitudes(p, 1)
    assert np.max(np.abs(A - Ao)) <= 1e-6 * np.max(np.abs(Ao))


@pytest.fixture(scope="module")
def full_pair():
    """FP64 and mixed handles on the benchmark workload (data and background drawn from
    the model)."""
    from ctpwa_amd import _native
    spec = W.WORKLOADS["jpsi_3pi"]
    s = W.generate_samples("jpsi_3pi", W.device_intensity, n_data=spec["n_data"] + spec["n_bkg"], n_bkg=0)
    ph, da, bk = s["phsp"], s["data"][:spec["n_data"]], s["data"][spec["n_data"]:]
    plan = plan_of("jpsi_3pi")
    wb = np.full(bk.shape[0], 0.8)
    print("case full size fp64 + mixed", flush=True)
    h64 = _native.Handle(plan.to_text(), ph, da, bk, None, None, wb)
    h32 = _native.Handle(plan.to_text(), ph, da, bk, None, None, wb, precision="mixed")
    return plan, h64, h32


def test_mixed_full_size_delta_nll(full_pair):
    plan, h64, h32 = full_pair
    base = W.nominal_params(plan, 1)
    for k, p in enumerate([base, perturbed(base, plan, 21), perturbed(base, plan, 22, 0.2)]):
        f64, g64, _ = _ng(h64, p)
        f32, g32, _ = _ng(h32, p)
        print("point", k, f64, f32, f32 - f64, rel(g32, g64), flush=True)
        assert abs(f32 - f64) <= 1e-4
        assert rel(g32, g64) <= 1e-5


def test_mixed_fit_shift(oracle_lib):
    """Fits of the same sample in both forms agree to 0.01 sigma."""
    import ctpwa_amd as ctpwa
    from ctpwa_amd.fit import Fitter
    spec = W.WORKLOADS["jpsi_3pi"]
    s = W.generate_samples("jpsi_3pi", W.device_intensity, n_data=100_000, n_phsp=500_000, n_bkg=0)
    print("case fit shift", flush=True)
    import os
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        cfg = os.path.join(d, "config.yml")
        text = "\n".join(l for l in spec["config"].splitlines()
                         if not l.strip().startswith(("bkg:", "bkg_weights:"))) + "\n"
        with open(cfg, "w") as f:
            f.write(text)
        samples_ = {"phsp": s["phsp"], "data": s["data"]}
        a64 = ctpwa.analysis(cfg, samples=samples_)
        a32 = ctpwa.analysis(cfg, samples=samples_, precision="mixed")
    plan = a64.plan
    p0 = W.nominal_params(plan, spec["seed"])
    n = plan.n_free
    v = p0[:n] + 1j * p0[n:2 * n]
    v = v / v[0]                                  # the driver fixes the first coupling to 1
    p0[:n], p0[n:2 * n] = v.real, v.imag
    p0 = perturbed(p0, plan, 31, 0.02)
    r64 = Fitter(a64).minimize(p0)
    r32 = Fitter(a32).minimize(p0)
    print("fits", r64.nll, r32.nll, r64.n_evals, r32.n_evals, r64.pos_def, r64.unconstrained, flush=True)
    # f0(980) has no allowed wave in J/psi -> pi0 f0: its mass and width do not enter
    th = [t.name for t in plan.theta]
    assert r64.unconstrained == [2 * n + th.index("f0_980_mass"), 2 * n + th.index("f0_980_width")]
    assert r64.pos_def
    free = r64.errors > 0
    shift = np.abs(r32.params - r64.params)[free] / r64.errors[free]
    print("max shift / sigma", shift.max(), flush=True)
    assert shift.max() <= 0.01
    assert abs(r32.nll - r64.nll) <= 1e-4 * 0.2 + 1e-4
