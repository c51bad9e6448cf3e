"""Table lineshapes (Hist, Interp) on the GPU against the oracle.

J/psi -> pi+ pi- pi0 with an Interp shape on an even grid (complex, spin 1) next to a
BWR, an Interp shape on an uneven grid (spin 2), a Hist with empty bins (fixed) and a
Hist whose optional mass parameter is free (it enters the parent's q0).  The init kernel
evaluates the tables once per event; the fused kernel multiplies the stored values in.
Bars: amplitudes 1e-10, NLL 1e-9, gradient 1e-7, Hessian 1e-6 (relative, FP64 form);
the mixed form keeps the table values in FP64 and is held to the mixed-form bars
(NLL 1e-9, gradient and Hessian 1e-5, amplitudes 1e-6 of the largest).
"""
import os

import numpy as np
import pytest
import torch

from helpers import perturbed, plan_of, rel, samples
from ctpwa_amd import workloads as W

pytestmark = pytest.mark.gpu


def _write_tables(d):
    rng = np.random.default_rng(0)
    x = np.linspace(0.25, 3.0, 56)
    z = 1.0 / (1.1 ** 2 - x ** 2 - 1j * 1.1 * 0.3) + 0.2 * np.exp(1j * x)
    np.savetxt(os.path.join(d, "interp_even.txt"), np.stack([x, z.real, z.imag], 1), fmt="%.17g")
    xu = np.sort(rng.uniform(0.25, 3.0, 40))
    xu[0], xu[-1] = 0.25, 3.0
    zu = np.exp(-((xu - 1.3) / 0.4) ** 2) * np.exp(2j * xu)
    np.savetxt(os.path.join(d, "interp_uneven.txt"), np.stack([xu, zu.real, zu.imag], 1), fmt="%.17g")
    v = np.exp(-((np.linspace(0.3, 2.9, 30) - 0.8) / 0.3) ** 2)
    v[:3] = 0.0
    np.savetxt(os.path.join(d, "hist.txt"), v, fmt="%.17g")


def _config(d):
    return """\
Particles:
  Jpsi: {J: 1, P: -1, mass: 3.0969, polarization: [1, -1]}
  pip:  {J: 0, P: -1, mass: 0.13957}
  pim:  {J: 0, P: -1, mass: 0.13957}
  pi0:  {J: 0, P: -1, mass: 0.13498}
Data:
  order: [pip, pim, pi0]
  data: [npy, data.npy]
  phsp: [npy, phsp.npy]
DecayChains:
  c:
    Jpsi:
      - [pi0, R_pm]
      - [pip, R_m0]
      - [pim, R_p0]
    R_pm: [pip, pim]
    R_m0: [pim, pi0]
    R_p0: [pip, pi0]
    intermediates:
      R_pm:
        - [J: 1, P: -1]: [X_even, rho770]
        - [J: 2, P: 1]: [T_uneven]
      R_m0:
        - [J: 1, P: -1]: [H_rho]
      R_p0:
        - [J: 1, P: -1]: [H_rho2]
Resonances:
  rho770:   {J: 1, P: -1, model: BWR, parameters: [0.775, 0.149], free: [0, 1]}
  X_even:   {J: 1, P: -1, model: Interp, file: %(d)s/interp_even.txt}
  T_uneven: {J: 2, P: 1, model: Interp, file: %(d)s/interp_uneven.txt}
  H_rho:    {J: 1, P: -1, model: Hist, file: %(d)s/hist.txt, bins: 30, range: [0.3, 2.9]}
  H_rho2:   {J: 1, P: -1, model: Hist, parameters: [0.8], free: [0], file: %(d)s/hist.txt, bins: 30,
             range: [0.3, 2.9]}
""" % {"d": d}


@pytest.fixture(scope="module", params=["fp64", "mixed"])
def case(request, tmp_path_factory, oracle_lib):
    from ctpwa_amd import _native
    d = str(tmp_path_factory.mktemp("tables"))
    _write_tables(d)
    plan = plan_of(_config(d))
    assert {r.model for b in plan.blocks for r in b.res} == {"BWR", "Interp", "Hist"}
    ph, da, _ = samples(20000, 4000, 0, 61)
    print("case tables", request.param, "20000 4000", flush=True)
    h = _native.Handle(plan.to_text(), ph, da, None, precision=request.param)
    O = oracle_lib.Oracle(plan.to_text(), 3, ph, da)
    yield request.param, plan, h, O
    h.close()


def test_table_amplitudes(case):
    prec, plan, h, O = case
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
        print(prec, "amp err", cls, err.max(), flush=True)
        if prec == "fp64":
            assert err.max() <= 1e-10
        else:
            assert np.max(np.abs(A - Ao)) <= 1e-6 * np.max(np.abs(Ao))


def test_table_nll_grad_hessian(case):
    prec, plan, h, O = case
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
        print(prec, k, nll_o, o[0], rel(o[1:], g_o), rel(H_g, H_o), flush=True)
        # the free Hist mass acts through the parent's barrier only
        assert abs(g_o[-1]) > 0.0
        if prec == "fp64":
            assert abs(o[0] - nll_o) <= 1e-9 * abs(nll_o)
            assert rel(o[1:], g_o) <= 1e-7
            assert rel(H_g, H_o) <= 1e-6
        else:
            assert abs(o[0] - nll_o) <= 1e-9 * abs(nll_o)
            assert rel(o[1:], g_o) <= 1e-5
            assert rel(H_g, H_o) <= 1e-5
