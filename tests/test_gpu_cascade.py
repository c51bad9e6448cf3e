"""Four-body cascades on the GPU against the oracle (J/psi -> K+ K- pi+ pi-).

Covers what the three-body workloads leave out: three-node chains whose middle node
carries a massive Wigner rotation (K* spin in the K1 frame), blocks holding two
resonances with free parameters (the K1 q0 depends on the K* mass parameter), Flatte
with two channels, an unfree resonance (phi) and spin-2 at both vertices.
Bars as for the three-body workloads: NLL 1e-9, gradient 1e-7, amplitudes 1e-10,
Hessian 1e-6 (relative).
"""
import numpy as np
import pytest
import torch

from helpers import perturbed, plan_of, rel, samples
from ctpwa_amd import workloads as W

pytestmark = pytest.mark.gpu

MASSES = W.WORKLOADS["jpsi_kkpipi"]["masses"]


@pytest.fixture(scope="module")
def cascade(oracle_lib):
    from ctpwa_amd import _native
    plan = plan_of("jpsi_kkpipi")
    ph, da, _ = samples(20000, 4000, 0, 71, masses=MASSES)
    print("case jpsi_kkpipi 20000 4000", flush=True)
    h = _native.Handle(plan.to_text(), ph, da, None)
    O = oracle_lib.Oracle(plan.to_text(), 4, ph, da)
    return plan, h, O


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


def test_cascade_amplitudes(cascade):
    plan, h, O = cascade
    p = perturbed(W.nominal_params(plan, 2), plan, 4)
    t = torch.tensor(p, dtype=torch.float64, device="cuda")
    for cls in (0, 1):
        n = h.n_events[cls]
        out = torch.empty((n, plan.npol, plan.n_amp, 2), dtype=torch.float64, device="cuda")
        h.amplitudes(t.data_ptr(), cls, out.data_ptr(), torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        A = out.cpu().numpy()
        A = A[..., 0] + 1j * A[..., 1]
        Ao = O.amplitudes(p, cls)
        err = np.max(np.abs(A - Ao), axis=(0, 1)) / np.max(np.abs(Ao), axis=(0, 1))
        print("amp err per amplitude", cls, err.max(), flush=True)
        assert err.max() <= 1e-10


def test_cascade_nll_grad_hessian(cascade):
    plan, h, O = cascade
    base = W.nominal_params(plan, 2)
    for k, p in enumerate([base, perturbed(base, plan, 5, 0.1)]):
        nll_o, g_o, H_o = O.hessian(p)
        nll_g, g_g, H_g = _run(h, p, hess=True)
        print(k, nll_o, nll_g, rel(g_g, g_o), rel(H_g, H_o), flush=True)
        assert abs(nll_g - nll_o) <= 1e-9 * abs(nll_o)
        assert rel(g_g, g_o) <= 1e-7
        assert rel(H_g, H_o) <= 1e-6
