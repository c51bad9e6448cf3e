"""The ``analysis`` drop-in on the GPU: reference API behaviour and full-size properties.

* getNLL + autograd and getHessian through a config.yml with npy samples, against the
  oracle on the same events;
* error behaviour of the reference's TORCH_CHECKs (wrong dtype, device, shape);
* size-independent properties at the benchmark size (2.5M phase-space, 500k data,
  50k background events, data and background drawn from the model): invariance of the
  NLL under a common rephasing and a common rescaling of the couplings (exact identities of the likelihood: the normalisation
  cancels |c|^2), central finite differences of the GPU NLL against the GPU gradient,
  and of the GPU gradient against the GPU Hessian;
* edge cases: a single data event, no background, data weights of zero.
"""
import os

import numpy as np
import pytest
import torch

from helpers import perturbed, plan_of, rel, samples
from ctpwa_amd import workloads as W

pytestmark = pytest.mark.gpu


def _write_cfg(tmp_path, name, ph, da, bk):
    text = W.WORKLOADS[name]["config"]
    if bk is None:
        text = "\n".join(l for l in text.splitlines() if not l.strip().startswith(("bkg:", "bkg_weights:"))) + "\n"
    d = str(tmp_path)
    np.save(os.path.join(d, "phsp.npy"), ph)
    np.save(os.path.join(d, "data.npy"), da)
    if bk is not None:
        np.save(os.path.join(d, "bkg.npy"), bk)
    cfg = os.path.join(d, "config.yml")
    with open(cfg, "w") as f:
        f.write(text)
    return cfg


def test_analysis_api_matches_oracle(tmp_path, oracle_lib):
    import ctpwa_amd as ctpwa
    ph, da, bk = samples(30000, 6000, 1500, 51)
    cfg = _write_cfg(tmp_path, "jpsi_3pi", ph, da, bk)
    print("case analysis jpsi_3pi", flush=True)
    ana = ctpwa.analysis(cfg)
    assert ana.isValid()
    plan = ana.plan
    assert (ana.getNVector(), ana.getNFreeTheta(), ana.getNParams()) == (10, 10, 30)
    O = oracle_lib.Oracle(plan.to_text(), 3, ph, da, bk, w_bkg=np.full(bk.shape[0], 0.8))
    p0 = perturbed(W.nominal_params(plan, 1), plan, 12)
    p = torch.tensor(p0, dtype=torch.float64, device="cuda", requires_grad=True)
    nll = ana.getNLL(p)
    assert nll.dim() == 0 and nll.dtype == torch.float64
    (g,) = torch.autograd.grad(nll, p)
    nll_o, g_o, H_o = O.hessian(p0)
    assert abs(nll.item() - nll_o) <= 1e-9 * abs(nll_o)
    assert rel(g.cpu().numpy(), g_o) <= 1e-7
    # chain rule through autograd: d(3 NLL)/dp = 3 grad
    p2 = p.detach().clone().requires_grad_(True)
    (3.0 * ana.getNLL(p2)).backward()
    assert rel(p2.grad.cpu().numpy(), 3.0 * g_o) <= 1e-7
    H = ana.getHessian(p.detach())
    assert H.shape == (30, 30) and H.device.type == "cuda"
    assert rel(H.cpu().numpy(), H_o) <= 1e-6
    # the reference's checks on the parameter tensor
    with pytest.raises(RuntimeError):
        ana.getNLL(p.detach().float())
    with pytest.raises(RuntimeError):
        ana.getNLL(p.detach().cpu())
    with pytest.raises(RuntimeError):
        ana.getNLL(p.detach()[:-1])


def test_getnll_is_the_c_abi_result_on_the_current_stream(tmp_path):
    """getNLL (the C++ autograd function) returns exactly what ctpwa_nll_grad writes, for
    a non-contiguous parameter view and on a side stream made current; its backward
    scales the gradient by the incoming gradient."""
    import ctpwa_amd as ctpwa
    ph, da, bk = samples(20000, 4000, 1000, 53)
    cfg = _write_cfg(tmp_path, "jpsi_3pi", ph, da, bk)
    print("case getNLL vs C ABI", flush=True)
    ana = ctpwa.analysis(cfg)
    n = ana.getNParams()
    p0 = perturbed(W.nominal_params(ana.plan, 1), ana.plan, 14)
    t = torch.tensor(p0, dtype=torch.float64, device="cuda")
    out = torch.empty(1 + n, dtype=torch.float64, device="cuda")
    ana._h.nll_grad(t.data_ptr(), out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    ref = out.cpu().numpy()
    big = torch.zeros(2 * n, dtype=torch.float64, device="cuda")
    big[::2] = t
    big.requires_grad_(True)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        nll = ana.getNLL(big[::2])
        (g,) = torch.autograd.grad(2.5 * nll, big)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    assert nll.item() == ref[0]
    gg = g.cpu().numpy()
    assert np.array_equal(gg[::2], 2.5 * ref[1:])
    assert np.all(gg[1::2] == 0.0)


def test_edge_samples(oracle_lib):
    """One data event, no background; a data sample whose weights are all zero."""
    from ctpwa_amd import _native
    plan = plan_of("jpsi_rho")
    ph, da, _ = samples(300, 1, 0, 53)
    print("case single data event", flush=True)
    h = _native.Handle(plan.to_text(), ph, da, None)
    O = oracle_lib.Oracle(plan.to_text(), 3, ph, da)
    p = perturbed(W.nominal_params(plan, None), plan, 13)
    t = torch.tensor(p, dtype=torch.float64, device="cuda")
    out = torch.empty(1 + h.n_params, dtype=torch.float64, device="cuda")
    h.nll_grad(t.data_ptr(), out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    nll_o, g_o = O.nll_grad(p)
    o = out.cpu().numpy()
    assert abs(o[0] - nll_o) <= 1e-9 * abs(nll_o)
    assert rel(o[1:], g_o) <= 1e-7
    # all-zero data weights: N_eff = 0, the likelihood is identically zero
    ph, da, _ = samples(2000, 500, 0, 55)
    print("case zero weights", flush=True)
    h = _native.Handle(plan.to_text(), ph, da, None, None, np.zeros(500), None)
    h.nll_grad(t.data_ptr(), out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    assert o[0] == 0.0 and np.all(o[1:] == 0.0)


@pytest.fixture(scope="module")
def full_size():
    """The benchmark workload at full size: 2.5M phase-space events, 500k data events by
    accept-reject on the model intensity, and 50k background events (weight 0.8) drawn
    the same way.  Events drawn from the model keep away from zeros of the intensity,
    where ln I has unbounded higher derivatives and central differences would measure
    their truncation error rather than the gradient (a phase-space-distributed sample
    of this size holds events with I ~ 1e-6 of the mean)."""
    from ctpwa_amd import _native
    spec = W.WORKLOADS["jpsi_3pi"]
    s = W.generate_samples("jpsi_3pi", W.device_intensity, n_data=spec["n_data"] + spec["n_bkg"], n_bkg=0)
    ph, da, bk = s["phsp"], s["data"][:spec["n_data"]], s["data"][spec["n_data"]:]
    plan = plan_of("jpsi_3pi")
    print("case full size", ph.shape[0], da.shape[0], bk.shape[0], flush=True)
    h = _native.Handle(plan.to_text(), ph, da, bk, None, None, np.full(bk.shape[0], 0.8))
    return plan, h


def _ng(h, p):
    t = torch.tensor(p, dtype=torch.float64, device="cuda")
    out = torch.empty(1 + h.n_params, dtype=torch.float64, device="cuda")
    h.nll_grad(t.data_ptr(), out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    return o[0], o[1:]


def test_full_size_invariances(full_size):
    plan, h = full_size
    n = plan.n_free
    p = perturbed(W.nominal_params(plan, 1), plan, 17)
    nll0, g0 = _ng(h, p)
    v = p[:n] + 1j * p[n:2 * n]
    for c in (np.exp(0.7j), 2.5, 0.3 * np.exp(-2.1j)):
        q = p.copy()
        w = v * c
        q[:n], q[n:2 * n] = w.real, w.imag
        nll1, g1 = _ng(h, q)
        print("invariance", c, nll0, nll1, flush=True)
        assert abs(nll1 - nll0) <= 1e-11 * abs(nll0)
        # theta gradient is unchanged, coupling gradient transforms as dL/dv* -> dL/dv* / conj(c)
        assert rel(g1[2 * n:], g0[2 * n:]) <= 1e-8
        gv0 = g0[:n] + 1j * g0[n:2 * n]
        gv1 = g1[:n] + 1j * g1[n:2 * n]
        assert np.max(np.abs(gv1 * np.conj(c) - gv0)) <= 1e-8 * np.max(np.abs(gv0))


def test_full_size_finite_differences(full_size):
    plan, h = full_size
    p = perturbed(W.nominal_params(plan, 1), plan, 19)
    _, g = _ng(h, p)
    t = torch.tensor(p, dtype=torch.float64, device="cuda")
    H = torch.empty((h.n_params, h.n_params), dtype=torch.float64, device="cuda")
    h.hessian(t.data_ptr(), H.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    H = H.cpu().numpy()
    assert np.max(np.abs(H - H.T)) <= 1e-9 * np.max(np.abs(H))
    for j in (0, 3, plan.n_free + 2, 2 * plan.n_free, 2 * plan.n_free + 5, h.n_params - 1):
        step = 1e-5 * max(abs(p[j]), 1e-2)
        pp, pm = p.copy(), p.copy()
        pp[j] += step
        pm[j] -= step
        fp, gp = _ng(h, pp)
        fm, gm = _ng(h, pm)
        fd = (fp - fm) / (2 * step)
        fdh = (gp - gm) / (2 * step)
        print("fd", j, g[j], fd, flush=True)
        assert abs(fd - g[j]) <= 1e-5 * max(abs(g[j]), 1e-3 * np.max(np.abs(g)))
        assert np.max(np.abs(fdh - H[j])) <= 1e-4 * np.max(np.abs(H[j]))


@pytest.mark.parametrize("precision,per_event", [("fp64", 4 * 16 + 7 * 8), ("mixed", 4 * 8 + 3 * 8 + 4 * 4)])
def test_mirror_polarisations_are_folded(tmp_path, oracle_lib, precision, per_event):
    """J/psi (M = +1, -1) into spin-0 particles: every SL tensor's M = -1 polarisation is
    the signed conjugate of its M = +1 one, so the handle stores one polarisation per SL
    tensor (4 tensors x 16 B (FP64) or 8 B (FP32) per event) and the kernels derive the
    other; the likelihood, gradient and the generated source agree with that."""
    import ctpwa_amd as ctpwa
    ph, da, bk = samples(9000, 2000, 500, 57)
    cfg = _write_cfg(tmp_path, "jpsi_3pi", ph, da, bk)
    print("case folded polarisations", precision, flush=True)
    ana = ctpwa.analysis(cfg, precision=precision)
    lay = ana.layout()
    assert lay["bytes_per_eval"] == sum(lay["n_events"]) * per_event
    assert lay["n_pol_stored"] == 1
    src = ana._h.kernel_source()
    assert "s0_1 = make_double2(" in src and "sl[1LL * ld" in src and "sl[4LL * ld" not in src
    plan = ana.plan
    O = oracle_lib.Oracle(plan.to_text(), 3, ph, da, bk, w_bkg=np.full(bk.shape[0], 0.8))
    p = perturbed(W.nominal_params(plan, 1), plan, 31)
    t = torch.tensor(p, device="cuda", requires_grad=True)
    nll = ana.getNLL(t)
    (g,) = torch.autograd.grad(nll, t)
    nll_o, g_o = O.nll_grad(p)
    tol = (1e-9, 1e-7) if precision == "fp64" else (1e-9, 1e-5)
    assert abs(float(nll.item()) - nll_o) <= tol[0] * abs(nll_o)
    assert rel(g.cpu().numpy(), g_o) <= tol[1]
