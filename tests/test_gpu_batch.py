"""Batched evaluation of many likelihoods of one model (ctpwa_batch_nll_grad).

Each likelihood's NLL and gradient from the batch must equal its own single evaluation
bit for bit (same event partition, same summation order), whatever the other members of
the batch are; handles of different models or precisions cannot share a batch.
"""
import numpy as np
import pytest
import torch

from helpers import perturbed, plan_of, samples
from ctpwa_amd import workloads as W

pytestmark = pytest.mark.gpu


def _solo(h, p):
    t = torch.tensor(p, dtype=torch.float64, device="cuda")
    out = torch.empty(1 + h.n_params, dtype=torch.float64, device="cuda")
    h.nll_grad(t.data_ptr(), out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_equals_solo(precision):
    from ctpwa_amd import _native
    plan = plan_of("jpsi_3pi")
    sizes = [(20000, 4000, 1000), (3000, 600, 0), (120000, 25000, 2500), (500, 100, 0), (45000, 9000, 0)]
    hs, ps = [], []
    print("case batch", precision, flush=True)
    for k, (n_p, n_d, n_b) in enumerate(sizes):
        ph, da, bk = samples(n_p, n_d, n_b, 90 + 3 * k)
        wb = None if bk is None else np.full(bk.shape[0], 0.8)
        hs.append(_native.Handle(plan.to_text(), ph, da, bk, None, None, wb, precision=precision))
        ps.append(perturbed(W.nominal_params(plan, 1), plan, 40 + k))
    solo = [_solo(h, p) for h, p in zip(hs, ps)]
    b = _native.Batch(8)
    P = torch.tensor(np.stack(ps), dtype=torch.float64, device="cuda")
    O = torch.zeros((len(hs), 1 + hs[0].n_params), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    for rep in range(2):
        O.zero_()
        b.nll_grad(hs, [P[j].data_ptr() for j in range(len(hs))], [O[j].data_ptr() for j in range(len(hs))], stream)
        torch.cuda.synchronize()
        got = O.cpu().numpy()
        for j in range(len(hs)):
            assert np.array_equal(got[j], solo[j]), (rep, j)
    # a subset in another order gives the same per-likelihood results
    order = [3, 0, 2]
    b.nll_grad([hs[j] for j in order], [P[j].data_ptr() for j in order], [O[j].data_ptr() for j in order], stream)
    torch.cuda.synchronize()
    got = O.cpu().numpy()
    for j in order:
        assert np.array_equal(got[j], solo[j])
    b.close()


def test_batch_rejects_mixed_models():
    from ctpwa_amd import _native
    ph, da, _ = samples(2000, 400, 0, 97)
    h1 = _native.Handle(plan_of("jpsi_3pi").to_text(), ph, da, None)
    h2 = _native.Handle(plan_of("jpsi_rho").to_text(), ph, da, None)
    h3 = _native.Handle(plan_of("jpsi_3pi").to_text(), ph, da, None, precision="mixed")
    b = _native.Batch(4)
    P = torch.zeros((3, 40), dtype=torch.float64, device="cuda")
    O = torch.zeros((3, 41), dtype=torch.float64, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    for other in (h2, h3):
        with pytest.raises(_native.NativeError):
            b.nll_grad([h1, other], [P[0].data_ptr(), P[1].data_ptr()], [O[0].data_ptr(), O[1].data_ptr()], s)
    with pytest.raises(ValueError):
        b.nll_grad([h1] * 5, [P[0].data_ptr()] * 5, [O[0].data_ptr()] * 5, s)
    b.close()


GRAPH_OFF = ("graph replays of batched evaluations are off by default: a fit-service run through them faulted "
             "the GPU (the captured prep launch kept the first job's model-table pointers, freed when that fit "
             "retired); the fix, per-job prep arguments in the descriptors, has not run on the GPU yet")


@pytest.mark.parametrize("mode", ["eager", pytest.param("graph", marks=pytest.mark.skip(reason=GRAPH_OFF))])
def test_repeated_lists_new_points_and_replaced_handles(mode, monkeypatch):
    """The same job list called again with new points written into the (unchanged)
    parameter buffers, with single evaluations of the same handles run in between, returns
    the single evaluations' bits; so do changed job lists, the return to the first list, and
    a handle destroyed and another created in its place (it may take the same address) with
    the same parameter and output buffers.  mode "graph" (CTPWA_BATCH_GRAPH=1): a repeated
    list reuses the uploaded descriptors and replays the captured launches."""
    from ctpwa_amd import _native
    plan = plan_of("jpsi_3pi")
    sizes = [(20000, 4000, 1000), (3000, 600, 0), (60000, 12000, 0), (500, 100, 0), (9000, 1800, 0)]
    hs = []
    print("case batch repeated lists", mode, flush=True)
    for k, (n_p, n_d, n_b) in enumerate(sizes):
        ph, da, bk = samples(n_p, n_d, n_b, 120 + 3 * k)
        wb = None if bk is None else np.full(bk.shape[0], 0.8)
        hs.append(_native.Handle(plan.to_text(), ph, da, bk, None, None, wb, precision="mixed"))
    base = W.nominal_params(plan, 1)
    if mode == "graph":
        monkeypatch.setenv("CTPWA_BATCH_GRAPH", "1")
    b = _native.Batch(8)
    monkeypatch.delenv("CTPWA_BATCH_GRAPH", raising=False)
    P = torch.zeros((len(hs), hs[0].n_params), dtype=torch.float64, device="cuda")
    O = torch.zeros((len(hs), 1 + hs[0].n_params), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    lists = [list(range(5)), [4, 1, 2], [3], list(range(5)), [4, 1, 2]]
    rep = 0
    for members in lists:
        for it in range(3):
            pts = [perturbed(base, plan, 1000 + 37 * rep + j) for j in range(len(hs))]
            P.copy_(torch.tensor(np.stack(pts), dtype=torch.float64))
            O.fill_(-1.0)
            b.nll_grad([hs[j] for j in members], [P[j].data_ptr() for j in members],
                       [O[j].data_ptr() for j in members], stream)
            torch.cuda.synchronize()
            got = O.cpu().numpy()
            for j in members:
                assert np.array_equal(got[j], _solo(hs[j], pts[j])), (members, it, j)
            rep += 1
    for k in (0, 3):
        hs[k].close()
        ph, da, _ = samples(7000 + 1000 * k, 1500, 0, 200 + k)
        hs[k] = _native.Handle(plan.to_text(), ph, da, None, precision="mixed")
        pts = [perturbed(base, plan, 3000 + 11 * k + j) for j in range(len(hs))]
        P.copy_(torch.tensor(np.stack(pts), dtype=torch.float64))
        b.nll_grad(hs, [P[j].data_ptr() for j in range(len(hs))], [O[j].data_ptr() for j in range(len(hs))], stream)
        torch.cuda.synchronize()
        got = O.cpu().numpy()
        for j in range(len(hs)):
            assert np.array_equal(got[j], _solo(hs[j], pts[j])), (k, j)
    b.close()
