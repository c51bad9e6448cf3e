"""Post-fit outputs on the GPU: the amplitude Gram (ctpwa_amplitude_gram) and per-event
partial-wave intensities (ctpwa_partial_intensities) against the oracle's amplitudes,
and getFitFractions / writeResult through the ``analysis`` drop-in."""
import os

import numpy as np
import pytest
import torch

from helpers import perturbed, plan_of, samples
from test_gpu_api import _write_cfg
from ctpwa_amd import results as R
from ctpwa_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_amplitude_gram_and_partial_intensities(precision, oracle_lib):
    from ctpwa_amd import _native
    plan = plan_of("jpsi_3pi")
    ph, da, bk = samples(20000, 4000, 1000, 81)
    wb = np.full(bk.shape[0], 0.8)
    print("case results", precision, flush=True)
    h = _native.Handle(plan.to_text(), ph, da, bk, None, None, wb, precision=precision)
    O = oracle_lib.Oracle(plan.to_text(), 3, ph, da, bk, w_bkg=wb)
    p = perturbed(W.nominal_params(plan, 1), plan, 17)
    t = torch.tensor(p, dtype=torch.float64, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    tol = 1e-10 if precision == "fp64" else 1e-6
    for cls, n in ((0, 20000), (1, 4000), (2, 1000)):
        A = O.amplitudes(p, cls)
        Qo = np.einsum("epa,epb->ab", np.conj(A), A)
        G = torch.empty((plan.n_amp, plan.n_amp, 2), dtype=torch.float64, device="cuda")
        h.amplitude_gram(t.data_ptr(), cls, G.data_ptr(), s)
        torch.cuda.synchronize()
        Q = G.cpu().numpy()
        Q = Q[..., 0] + 1j * Q[..., 1]
        assert np.max(np.abs(Q - Qo)) <= tol * np.max(np.abs(Qo)), cls
        nb = len(plan.blocks)
        out = torch.empty((nb + 1, n), dtype=torch.float64, device="cuda")
        h.partial_intensities(t.data_ptr(), cls, out.data_ptr(), s)
        torch.cuda.synchronize()
        o = out.cpu().numpy()
        v = R.coupling_vector(plan, p)
        for b, B in enumerate(plan.blocks):
            Pb = np.einsum("epa,a->ep", A[:, :, B.site:B.site + B.nsl], v[B.site:B.site + B.nsl])
            ref = np.sum(np.abs(Pb) ** 2, axis=1)
            assert np.max(np.abs(o[b] - ref)) <= 10 * tol * max(1e-30, np.max(np.abs(o[-1]))), (cls, b)
        I = O.intensity(p, cls)
        assert np.max(np.abs(o[-1] - I)) <= 10 * tol * np.max(I)


def test_fit_fractions_and_write_result(tmp_path, oracle_lib):
    import ctpwa_amd as ctpwa
    ph, da, bk = samples(20000, 4000, 1000, 83)
    truth = samples(30000, 1, 0, 84)[0]
    cfg = _write_cfg(tmp_path, "jpsi_3pi", ph, da, bk)
    print("case fit fractions", flush=True)
    ana = ctpwa.analysis(cfg, samples={"phsp_truth": truth})
    plan = ana.plan
    p = perturbed(W.nominal_params(plan, 1), plan, 19)
    pt = torch.tensor(p, dtype=torch.float64, device="cuda")
    # the complex-vector form takes theta from the last evaluated point
    ana.getNLL(pt)
    n = plan.n_free
    vec = torch.complex(pt[:n], pt[n:2 * n])
    rng = np.random.default_rng(3)
    B = rng.standard_normal((ana.getNParams(), ana.getNParams()))
    Hpd = torch.tensor(B @ B.T + ana.getNParams() * np.eye(ana.getNParams()), dtype=torch.float64, device="cuda")
    res = ana.getFitFractions(vec, Hpd)
    assert res.shape == (len(plan.blocks), 2) and res.dtype == torch.float64 and not res.is_cuda
    Ot = oracle_lib.Oracle(plan.to_text(), 3, truth, truth[:1], None)
    A = Ot.amplitudes(p, 0)
    Q = np.einsum("epa,epb->ab", np.conj(A), A)
    ff = R.fit_fractions(Q, plan, p)
    err = R.fit_fraction_errors(R.fit_fraction_jacobian(Q, plan, p), Hpd.cpu().numpy(), n)
    r = res.numpy()
    assert np.allclose(r[:, 0], ff, rtol=1e-9, atol=1e-14)
    assert np.allclose(r[:, 1], err, rtol=1e-8, atol=1e-14)
    assert abs(r[:, 0].sum() - 1.0) < 1e-12
    # the whole-parameter form gives the same
    assert np.allclose(ana.getFitFractions(pt, Hpd).numpy(), r, rtol=1e-12, atol=0)
    with pytest.raises(RuntimeError):
        ana.getFitFractions(vec.to(torch.complex64), Hpd)
    # writeResult: interference matrix and per-event weights over the phase-space sample
    path = ana.writeResult(pt, os.path.join(str(tmp_path), "result.root"), is_saved_weight=1)
    assert path.endswith(".npz")
    z = np.load(path, allow_pickle=False)
    Op = oracle_lib.Oracle(plan.to_text(), 3, ph, da, bk, w_bkg=np.full(bk.shape[0], 0.8))
    Ap = Op.amplitudes(p, 0)
    Qp = np.einsum("epa,epb->ab", np.conj(Ap), Ap)
    X, total = R.interference_matrix(Qp, plan, p)
    assert np.allclose(z["interference"], X, rtol=1e-9, atol=1e-13)
    assert abs(float(z["phsp_integral"]) - total) <= 1e-9 * total
    assert float(z["data_integral"]) == 4000 - 0.8 * 1000
    tw = z["totalweight"]
    assert abs(tw.sum() - float(z["data_integral"])) <= 1e-9 * float(z["data_integral"])
    nb = len(plan.blocks)
    parts = sum(z["weight_%d" % k].sum() for k in range(nb))
    assert abs(parts / tw.sum() - np.trace(X)) <= 1e-9
    assert list(z["legends"]) == R.partial_names(plan)
    assert np.array_equal(z["%s_E" % plan.final_order[0]], ph[:, 0, 0])
    assert np.array_equal(z["%s_pz" % plan.final_order[2]], ph[:, 2, 3])
    # per-event pair terms: weights are conserved event by event, and their sums are the
    # off-diagonal interference entries
    pairs = sum(z["interf_%d_%d" % (i, j)] for i in range(nb) for j in range(i + 1, nb))
    wsum = sum(z["weight_%d" % k] for k in range(nb))
    assert np.abs(wsum + pairs - tw).max() <= 1e-10 * tw.max()
    assert abs(z["interf_0_1"].sum() / tw.sum() - X[0, 1]) <= 1e-9
    # coherent partial-wave amplitudes against the oracle's amplitudes and couplings
    v = R.coupling_vector(plan, p)
    Pw = np.stack([Ap[:, :, B.site:B.site + B.nsl] @ v[B.site:B.site + B.nsl] for B in plan.blocks])
    Pg = ana.getPartialAmplitudes(pt, 0).cpu().numpy()
    assert Pg.shape == Pw.shape and np.abs(Pg - Pw).max() <= 1e-10 * np.abs(Pw).max()
    # a subset of waves: per-event total sum_pol |P_0 + P_2|^2, integral over it
    path = ana.writeResult(pt, os.path.join(str(tmp_path), "subset"), is_saved_weight=1, waves=[0, 2])
    zs = np.load(path, allow_pickle=False)
    ev = np.sum(np.abs(Pw[0] + Pw[2]) ** 2, axis=1)
    assert abs(float(zs["phsp_integral"]) - ev.sum()) <= 1e-9 * ev.sum()
    di = float(zs["data_integral"])
    assert np.abs(zs["totalweight"] - ev * di / ev.sum()).max() <= 1e-9 * zs["totalweight"].max()
    assert np.allclose(zs["interference"], X, rtol=1e-9, atol=1e-13)
    with pytest.raises(RuntimeError):
        ana.writeResult(pt, os.path.join(str(tmp_path), "bad"), waves=[nb])


def test_fit_fractions_need_a_truth_sample(tmp_path):
    import ctpwa_amd as ctpwa
    ph, da, _ = samples(2000, 400, 0, 85)
    ana = ctpwa.analysis(_write_cfg(tmp_path, "jpsi_3pi", ph, da, None))
    pt = torch.tensor(W.nominal_params(ana.plan, 1), dtype=torch.float64, device="cuda")
    with pytest.raises(RuntimeError):
        ana.getFitFractions(pt)
