"""Identical-particle symmetrisation on the GPU against the oracle.

The three configurations of the CPU tests (J/psi -> gamma pi0 pi0 through f0/f2 and
omega, a fermion pair J/psi -> n n pi0 with the minus sign of the exchange, three
identical pi0 with all five non-identity permutations) run through the device: amplitudes
of every permuted copy, NLL, gradient and Hessian (1e-10, 1e-9, 1e-7, 1e-6 relative), and
the per-event partial-wave intensities, which sum each wave's copies coherently.
"""
import numpy as np
import pytest
import torch

from helpers import perturbed, plan_of, rel
from test_identical_cpu import CASES, _sample
from ctpwa_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request, oracle_lib):
    from ctpwa_amd import _native
    text, masses, _, _ = CASES[request.param]
    plan = plan_of(text)
    ph = _sample(20000, 91, masses)
    da = _sample(4000, 92, masses)
    print("case", request.param, "20000 4000", flush=True)
    h = _native.Handle(plan.to_text(), ph, da, None)
    O = oracle_lib.Oracle(plan.to_text(), 3, ph, da)
    yield request.param, plan, h, O
    h.close()


def test_identical_amplitudes_and_partial_waves(case):
    name, plan, h, O = case
    p = perturbed(W.nominal_params(plan, 3), plan, 4)
    t = torch.tensor(p, dtype=torch.float64, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    n = plan.n_free
    v = np.array([plan.amp_ratio[a] * (p[plan.amp_free[a]] + 1j * p[n + plan.amp_free[a]])
                  for a in range(plan.n_amp)])
    wave = plan.wave_of()
    for cls in (0, 1):
        ne = h.n_events[cls]
        out = torch.empty((ne, plan.npol, plan.n_amp, 2), dtype=torch.float64, device="cuda")
        h.amplitudes(t.data_ptr(), cls, out.data_ptr(), s)
        pw = torch.empty((len(plan.waves) + 1, ne), dtype=torch.float64, device="cuda")
        h.partial_intensities(t.data_ptr(), cls, pw.data_ptr(), s)
        torch.cuda.synchronize()
        A = out.cpu().numpy()
        A = A[..., 0] + 1j * A[..., 1]
        Ao = O.amplitudes(p, cls)
        err = np.max(np.abs(A - Ao), axis=(0, 1)) / np.max(np.abs(Ao), axis=(0, 1))
        print(name, "amp err", cls, err.max(), flush=True)
        assert err.max() <= 1e-10
        # partial waves: a block and its permuted copies summed coherently
        P = np.zeros((ne, plan.npol, len(plan.waves)), dtype=complex)
        for b, B in enumerate(plan.blocks):
            P[:, :, wave[b]] += Ao[:, :, B.site:B.site + B.nsl] @ v[B.site:B.site + B.nsl]
        ref = np.concatenate([np.sum(np.abs(P) ** 2, axis=1).T,
                              np.sum(np.abs(P.sum(axis=2)) ** 2, axis=1)[None]], axis=0)
        got = pw.cpu().numpy()
        print(name, "partial err", cls, rel(got, ref), flush=True)
        assert rel(got, ref) <= 1e-10
        assert rel(got[-1], O.intensity(p, cls)) <= 1e-10


def test_identical_nll_grad_hessian(case):
    name, plan, h, O = case
    base = W.nominal_params(plan, 3)
    for k, p in enumerate([base, perturbed(base, plan, 5, 0.1)]):
        nll_o, g_o, H_o = O.hessian(p)
        t = torch.tensor(p, dtype=torch.float64, device="cuda")
        out = torch.empty(1 + h.n_params, dtype=torch.float64, device="cuda")
        Ht = torch.empty((h.n_params, h.n_params), dtype=torch.float64, device="cuda")
        s = torch.cuda.current_stream().cuda_stream
        h.nll_grad(t.data_ptr(), out.data_ptr(), s)
        h.hessian(t.data_ptr(), Ht.data_ptr(), s)
        torch.cuda.synchronize()
        o = out.cpu().numpy()
        H_g = Ht.cpu().numpy()
        print(name, k, nll_o, o[0], rel(o[1:], g_o), rel(H_g, H_o), flush=True)
        assert abs(o[0] - nll_o) <= 1e-9 * abs(nll_o)
        assert rel(o[1:], g_o) <= 1e-7
        assert rel(H_g, H_o) <= 1e-6
