"""HIP path against the CPU FP64 oracle on the same seeded events (through the C ABI).

Bars: NLL 1e-9 relative, gradient 1e-7 relative, amplitudes 1e-10 relative
(element-wise, normalised by the largest amplitude).
"""
import numpy as np
import pytest
import torch

from helpers import perturbed, plan_of, rel, samples
from ctpwa_amd import workloads as W

pytestmark = pytest.mark.gpu

CASES = [
    # name, n_phsp, n_data, n_bkg, w_bkg, seed
    ("jpsi_rho", 20000, 5000, 0, None, 11),
    ("jpsi_3pi", 30000, 6000, 1500, 0.8, 21),
]


def _handles(name, n_p, n_d, n_b, w_b, seed, oracle_lib):
    from ctpwa_amd import _native
    plan = plan_of(name)
    ph, da, bk = samples(n_p, n_d, n_b, seed)
    wb = None if bk is None else np.full(bk.shape[0], w_b)
    print("case", name, n_p, n_d, n_b, flush=True)
    h = _native.Handle(plan.to_text(), ph, da, bk, None, None, wb)
    O = oracle_lib.Oracle(plan.to_text(), 3, ph, da, bk, w_bkg=wb)
    return plan, h, O


def _gpu_nll_grad(h, p):
    t = torch.tensor(p, dtype=torch.float64, device="cuda")
    out = torch.empty(1 + h.n_params, dtype=torch.float64, device="cuda")
    h.nll_grad(t.data_ptr(), out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    return o[0], o[1:]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_nll_grad_parity(case, oracle_lib):
    from ctpwa_amd import workloads as W
    plan, h, O = _handles(*case, oracle_lib)
    base = W.nominal_params(plan, None if case[0] == "jpsi_rho" else 1)
    for k, p in enumerate([base, perturbed(base, plan, 5), perturbed(base, plan, 6, 0.2)]):
        nll_o, g_o = O.nll_grad(p)
        nll_g, g_g = _gpu_nll_grad(h, p)
        print(k, nll_o, nll_g, rel(nll_g, nll_o), rel(g_g, g_o), flush=True)
        assert abs(nll_g - nll_o) <= 1e-9 * abs(nll_o)
        assert rel(g_g, g_o) <= 1e-7


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_amplitude_parity(case, oracle_lib):
    from ctpwa_amd import workloads as W
    plan, h, O = _handles(*case, oracle_lib)
    p = perturbed(W.nominal_params(plan, 1), plan, 9)
    t = torch.tensor(p, dtype=torch.float64, device="cuda")
    for cls in (0, 1):
        n = h.n_events[cls]
        out = torch.empty((n, plan.npol, plan.n_amp, 2), dtype=torch.float64, device="cuda")
        h.amplitudes(t.data_ptr(), cls, out.data_ptr(), torch.cuda.current_stream().cuda_stream)
        inten = torch.empty(n, dtype=torch.float64, device="cuda")
        h.intensity(t.data_ptr(), cls, inten.data_ptr(), torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        A = out.cpu().numpy()
        A = A[..., 0] + 1j * A[..., 1]
        Ao = O.amplitudes(p, cls)
        assert np.max(np.abs(A - Ao)) <= 1e-10 * np.max(np.abs(Ao))
        Io = O.intensity(p, cls)
        assert rel(inten.cpu().numpy(), Io) <= 1e-10


def test_deterministic(oracle_lib):
    """Two evaluations at the same point are bit-identical (fixed reduction order)."""
    from ctpwa_amd import workloads as W
    plan, h, _ = _handles("jpsi_3pi", 30000, 6000, 1500, 0.8, 31, oracle_lib)
    p = W.nominal_params(plan, 1)
    a = _gpu_nll_grad(h, p)
    b = _gpu_nll_grad(h, p)
    assert a[0] == b[0]
    assert np.array_equal(a[1], b[1])


def test_constrained_plan_parity(oracle_lib):
    """trans [R_m0, R_p0]: -1 and var_equal [rho1450_width, rho1700_width]: shared
    couplings with ratio -1 and a shared resonance parameter, against the oracle."""
    from ctpwa_amd import _native
    text = (W_JPSI_3PI + "Constraints:\n  trans:\n    - [R_m0, R_p0]: -1\n"
            "  var_equal:\n    - [rho1450_width, rho1700_width]\n")
    plan = plan_of(text)
    ph, da, bk = samples(30000, 6000, 1500, 23)
    wb = np.full(bk.shape[0], 0.8)
    print("case constrained", flush=True)
    h = _native.Handle(plan.to_text(), ph, da, bk, None, None, wb)
    O = oracle_lib.Oracle(plan.to_text(), 3, ph, da, bk, w_bkg=wb)
    rng = np.random.default_rng(4)
    p = np.concatenate([rng.uniform(-1, 1, 2 * plan.n_free), [t.init for t in plan.theta]])
    nll_o, g_o, H_o = O.hessian(p)
    nll_g, g_g = _gpu_nll_grad(h, p)
    t = torch.tensor(p, dtype=torch.float64, device="cuda")
    H = torch.empty((h.n_params, h.n_params), dtype=torch.float64, device="cuda")
    h.hessian(t.data_ptr(), H.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert abs(nll_g - nll_o) <= 1e-9 * abs(nll_o)
    assert rel(g_g, g_o) <= 1e-7
    assert rel(H.cpu().numpy(), H_o) <= 1e-6


from ctpwa_amd.workloads import JPSI_3PI as W_JPSI_3PI  # noqa: E402


MODELS_3PI = W_JPSI_3PI.replace(
    "  rho770:  {J: 1, P: -1, model: BWR, parameters: [0.775, 0.149], free: [0, 1]}",
    "  rho770:  {J: 1, P: -1, model: GS, parameters: [0.775, 0.149], free: [0, 1], channels: [[0.13975, 0.13975]]}").replace(
    "  rho1450: {J: 1, P: -1, model: BWR, parameters: [1.465, 0.40],  free: [0, 1]}",
    "  rho1450: {J: 1, P: -1, model: BW, parameters: [1.465, 0.40],  free: [0, 1]}").replace(
    "  rho1700: {J: 1, P: -1, model: BWR, parameters: [1.72, 0.25],   free: [0, 1]}",
    "  rho1700: {J: 1, P: -1, model: ONE, parameters: [1.72]}") + "Constraints:\n  bf_d: 4.0\n"


def test_all_builtin_lineshapes_parity(oracle_lib):
    """GS, BW, ONE and BWR (Flatte is in the four-body cascade) with a non-default barrier
    radius, against the oracle: NLL, gradient, Hessian."""
    from ctpwa_amd import _native
    plan = plan_of(MODELS_3PI)
    models = sorted({r.model for b in plan.blocks for r in b.res})
    assert models == ["BW", "BWR", "GS", "ONE"]
    ph, da, bk = samples(30000, 6000, 1500, 25)
    wb = np.full(bk.shape[0], 0.8)
    print("case models", flush=True)
    h = _native.Handle(plan.to_text(), ph, da, bk, None, None, wb)
    O = oracle_lib.Oracle(plan.to_text(), 3, ph, da, bk, w_bkg=wb)
    p = perturbed(W.nominal_params(plan, 1), plan, 15)
    nll_o, g_o, H_o = O.hessian(p)
    nll_g, g_g = _gpu_nll_grad(h, p)
    t = torch.tensor(p, dtype=torch.float64, device="cuda")
    H = torch.empty((h.n_params, h.n_params), dtype=torch.float64, device="cuda")
    h.hessian(t.data_ptr(), H.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    print(nll_o, nll_g, rel(g_g, g_o), rel(H.cpu().numpy(), H_o), flush=True)
    assert abs(nll_g - nll_o) <= 1e-9 * abs(nll_o)
    assert rel(g_g, g_o) <= 1e-7
    assert rel(H.cpu().numpy(), H_o) <= 1e-6
