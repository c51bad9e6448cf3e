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
