"""ctpwa_nll_grad_noalloc (the NLL+gradient entry point that allocates nothing: its
partial sums live in a static device array) against ctpwa_nll_grad, through the C ABI:
bit-identical outputs on two calls in a row, the second on a smaller likelihood (fewer
events, fewer workgroups), and again on the first; a model whose partials exceed the
array is rejected with an error."""
import ctypes

import numpy as np
import pytest
import torch

from helpers import perturbed, plan_of, samples
from ctpwa_amd import _native
from ctpwa_amd import workloads as W

pytestmark = pytest.mark.gpu


def _eval(h, fn, p):
    dev = torch.device("cuda", 0)
    pt = torch.tensor(p, dtype=torch.float64, device=dev)
    out = torch.full((1 + p.size,), np.nan, dtype=torch.float64, device=dev)
    s = torch.cuda.current_stream(dev)
    _native.check(fn(h.h, ctypes.c_void_p(pt.data_ptr()), ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(s.cuda_stream)),
                  "evaluation")
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_noalloc_matches_nll_grad():
    L = _native.lib()
    plan = plan_of("jpsi_3pi")
    text = plan.to_text()
    big = samples(200_000, 40_000, 8_000, 71)
    small = samples(30_000, 6_000, 1_000, 72)
    print("case no-alloc: 248k then 37k events", flush=True)
    hb = _native.Handle(text, big[0], big[1], big[2], None, None, np.full(big[2].shape[0], 0.8), device=0)
    hs = _native.Handle(text, small[0], small[1], small[2], None, None, np.full(small[2].shape[0], 0.8), device=0)
    p0 = perturbed(W.nominal_params(plan, 1), plan, 3)
    p1 = perturbed(W.nominal_params(plan, 1), plan, 4)
    for h, p in ((hb, p0), (hs, p1), (hb, p1)):
        ref = _eval(h, L.ctpwa_nll_grad, p)
        got = _eval(h, L.ctpwa_nll_grad_noalloc, p)
        assert np.all(np.isfinite(ref))
        assert np.array_equal(got, ref), (got[:4], ref[:4])
    hb.close()
    hs.close()


def test_noalloc_rejects_oversized_partials():
    plan = plan_of("jpsi_kkpipi_wide")
    ph, da, _ = samples(1_000_000, 2_000, 0, 73, masses=W.WORKLOADS["jpsi_kkpipi_wide"]["masses"])
    print("case no-alloc: oversized partials", flush=True)
    h = _native.Handle(plan.to_text(), ph, da, None, device=0)
    p = W.nominal_params(plan, 4)
    with pytest.raises(_native.NativeError, match="scratch"):
        _eval(h, _native.lib().ctpwa_nll_grad_noalloc, p)
    h.close()
