"""Massless daughters on the GPU against the oracle.

The workloads have only massive final states, so the device's MasslessTrans branch
(src/Amplitude.cu:409-470, followed by the helicity rotation of src/Amplitude.cu:570-586)
runs only here:

* radiative J/psi -> K+ K- gamma: K*(892)+- -> K gamma with the photon as first and
  as second daughter of a moving K*, and J/psi -> gamma f0/f2 with f -> K+ K-;
* a massless spin-1/2 pair J/psi -> nu nubar pi0 through spin-1/2 and 3/2 states, the
  half-integer branch of the helicity phase.

The reference takes a daughter as massless when sqrt(max(E^2 - p^2, 0)) <= 1e-10; the
photon and neutrino energies are set to |p| (1 - 1e-12), so both the oracle and the
device take the massless branch whatever the rounding of E^2 - p^2.  Bars as for the
workloads: amplitudes 1e-10, NLL 1e-9, gradient 1e-7, Hessian 1e-6 (relative).
"""
import numpy as np
import pytest
import torch

from helpers import perturbed, plan_of, rel
from ctpwa_amd import workloads as W

pytestmark = pytest.mark.gpu

M_K = 0.493677

RADIATIVE = """\
Particles:
  Jpsi: {J: 1, P: -1, mass: 3.0969, polarization: [1, -1]}
  gam: {J: 1, P: -1, mass: 0.0}
  Kp:  {J: 0, P: -1, mass: 0.493677}
  Km:  {J: 0, P: -1, mass: 0.493677}
Data:
  order: [gam, Kp, Km]
  data: [npy, data.npy]
  phsp: [npy, phsp.npy]
DecayChains:
  kstm:
    Jpsi: [[Kp, Kstm]]
    Kstm: [Km, gam]
    intermediates:
      Kstm:
        - [J: 1, P: -1]: [Kst892]
  kstp:
    Jpsi: [[Km, Kstp]]
    Kstp: [gam, Kp]
    intermediates:
      Kstp:
        - [J: 1, P: -1]: [Kst892]
  gkk:
    Jpsi: [[gam, R_KK]]
    R_KK: [Kp, Km]
    intermediates:
      R_KK:
        - [J: 0, P: 1]: [f0_1500]
        - [J: 2, P: 1]: [f2_1525]
Resonances:
  Kst892:  {J: 1, P: -1, model: BWR, parameters: [0.8955, 0.0473], free: [0, 1]}
  f0_1500: {J: 0, P: 1, model: BWR, parameters: [1.506, 0.112], free: [0, 1]}
  f2_1525: {J: 2, P: 1, model: BWR, parameters: [1.517, 0.086], free: [0, 1]}
"""

NEUTRINOS = """\
Particles:
  Jpsi: {J: 1, P: -1, mass: 3.0969, polarization: [1, -1]}
  nu:  {J: 1/2, P: 1, mass: 0.0}
  nub: {J: 1/2, P: -1, mass: 0.0}
  pi0: {J: 0, P: -1, mass: 0.13495}
Data:
  order: [nu, nub, pi0]
  data: [npy, data.npy]
  phsp: [npy, phsp.npy]
DecayChains:
  n1:
    Jpsi: [[nub, N1]]
    N1: [nu, pi0]
    intermediates:
      N1:
        - [J: 1/2, P: 1]: [N1440]
        - [J: 3/2, P: -1]: [N1520]
  n2:
    Jpsi: [[N2, nu]]
    N2: [pi0, nub]
    intermediates:
      N2:
        - [J: 1/2, P: -1]: [N1440b]
Resonances:
  N1440:  {J: 1/2, P: 1, model: BWR, parameters: [1.44, 0.35], free: [0, 1]}
  N1520:  {J: 3/2, P: -1, model: BWR, parameters: [1.515, 0.11], free: [0, 1]}
  N1440b: {J: 1/2, P: -1, model: BWR, parameters: [1.44, 0.35], free: [0, 1]}
"""

CASES = {
    "radiative": (RADIATIVE, (0.0, M_K, M_K), (0,)),
    "neutrinos": (NEUTRINOS, (0.0, 0.0, 0.13495), (0, 1)),
}


def massless_sample(n, seed, masses, massless):
    ev = W.phase_space(n, seed=seed, masses=masses)
    for i in massless:
        ev[:, i, 0] = np.sqrt((ev[:, i, 1:] ** 2).sum(-1)) * (1.0 - 1e-12)
    return ev


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request, oracle_lib):
    from ctpwa_amd import _native
    text, masses, massless = CASES[request.param]
    plan = plan_of(text)
    ph = massless_sample(20000, 81, masses, massless)
    da = massless_sample(4000, 82, masses, massless)
    print("case", request.param, "20000 4000", flush=True)
    h = _native.Handle(plan.to_text(), ph, da, None)
    O = oracle_lib.Oracle(plan.to_text(), 3, ph, da)
    yield request.param, plan, h, O
    h.close()


def test_massless_amplitudes(case):
    name, plan, h, O = case
    p = perturbed(W.nominal_params(plan, 3), plan, 4)
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
        print(name, "amp err", cls, err.max(), flush=True)
        assert err.max() <= 1e-10


def test_massless_nll_grad_hessian(case):
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
