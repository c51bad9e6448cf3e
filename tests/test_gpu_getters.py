"""The reference's remaining ``analysis`` getters on the GPU (src/main.cu:91-131).

* getPhspTensor / getDataTensor / getBkgTensor: the stored amplitudes at the current
  theta without couplings, flattened [event][pol][amp] complex128, against the oracle's
  amplitudes at the initial theta, after reCalcAmp(theta) and after getNLL at a new point;
  with identical particles, against the brute-force sum over exchanged events of the
  unsymmetrised amplitudes;
* getBkgWeightsTensor, getSLVectors, getConstraintsIndex/Values, setFitMode(0).
"""
import numpy as np
import pytest
import torch

from helpers import perturbed, samples
from test_gpu_api import _write_cfg
from test_identical_cpu import BOSONS, CASES, _permuted, _sample, _strip
from ctpwa_amd import workloads as W

pytestmark = pytest.mark.gpu


def _flat(A):
    return torch.from_numpy(np.ascontiguousarray(A)).reshape(-1)


def test_amplitude_getters(tmp_path, oracle_lib):
    import ctpwa_amd as ctpwa
    ph, da, bk = samples(8000, 2000, 500, 33)
    cfg = _write_cfg(tmp_path, "jpsi_3pi", ph, da, bk)
    print("case getters jpsi_3pi", flush=True)
    ana = ctpwa.analysis(cfg)
    plan = ana.plan
    O = oracle_lib.Oracle(plan.to_text(), 3, ph, da, bk, w_bkg=np.full(bk.shape[0], 0.8))
    p0 = W.nominal_params(plan, 1)
    for get, cls in ((ana.getPhspTensor, 0), (ana.getDataTensor, 1), (ana.getBkgTensor, 2)):
        t = get()
        assert t.dtype == torch.complex128 and t.is_cuda and t.dim() == 1
        assert t.numel() == (ph, da, bk)[cls].shape[0] * plan.npol * plan.n_amp
        ref = _flat(O.amplitudes(p0, cls))
        assert (t.cpu() - ref).abs().max() <= 1e-10 * ref.abs().max()
    # reCalcAmp moves the stored amplitudes to a new theta; so does getNLL
    p1 = perturbed(p0, plan, 3, 0.1)
    ana.reCalcAmp(torch.tensor(p1[2 * plan.n_free:]))
    ref = _flat(O.amplitudes(p1, 1))
    assert (ana.getDataTensor().cpu() - ref).abs().max() <= 1e-10 * ref.abs().max()
    p2 = perturbed(p0, plan, 4, 0.1)
    ana.getNLL(torch.tensor(p2, device="cuda"))
    ref = _flat(O.amplitudes(p2, 0))
    assert (ana.getPhspTensor().cpu() - ref).abs().max() <= 1e-10 * ref.abs().max()
    w = ana.getBkgWeightsTensor()
    assert w.dtype == torch.float32 and w.numel() == bk.shape[0] and torch.all(w == 0.8)
    sv = ana.getSLVectors()
    assert sv.dtype == torch.int32 and sv.tolist() == [b.nsl for b in plan.blocks]
    assert ana.getConstraintsIndex() == [] and ana.getConstraintsValues() == []
    n = ana.getNVector()
    ana.setFitMode(0)
    assert ana.getFitMode() == 0 and ana.getNVector() == n
    with pytest.raises(RuntimeError):
        ana.setFitMode(2)
    with pytest.raises(RuntimeError):
        ana.getSLAmpsTensor()


def test_amplitude_getters_with_identical_particles(tmp_path, oracle_lib):
    import os
    import ctpwa_amd as ctpwa
    from helpers import plan_of
    _, masses, idx, _ = CASES["bosons"]
    ph, da = _sample(4000, 41, masses), _sample(1000, 42, masses)
    cfgp = os.path.join(str(tmp_path), "config.yml")
    with open(cfgp, "w") as f:
        f.write(BOSONS)
    print("case getters identical", flush=True)
    ana = ctpwa.analysis(cfgp, samples={"phsp": ph, "data": da})
    plan = ana.plan
    base = _strip(plan_of(BOSONS))
    assert ana.getAmplitudeNames() == base.amp_names
    assert ana.getSLVectors().tolist() == [b.nsl for b in base.blocks]
    p = W.nominal_params(base, 3)
    ref = 0.0
    for perm in ((0, 1), (1, 0)):
        ref = ref + oracle_lib.Oracle(base.to_text(), 3, ph, _permuted(da, idx, perm)).amplitudes(p, 1)
    t = ana.getDataTensor().cpu()
    assert t.numel() == da.shape[0] * plan.npol * base.n_amp
    assert (t - _flat(ref)).abs().max() <= 1e-10 * _flat(ref).abs().max()


def test_theta_of_last_evaluation_survives_other_calls(tmp_path, oracle_lib):
    """getNLL(A), then getIntensity at B: the stored amplitudes, and getFitFractions of a
    complex vector, stay at A's theta (the theta of the last getNLL / getHessian)."""
    import ctpwa_amd as ctpwa
    ph, da, bk = samples(6000, 1500, 0, 35)
    cfg = _write_cfg(tmp_path, "jpsi_3pi", ph, da, None)
    print("case theta of the last evaluation", flush=True)
    ana = ctpwa.analysis(cfg, samples={"phsp_truth": ph})
    plan = ana.plan
    O = oracle_lib.Oracle(plan.to_text(), 3, ph, da)
    p0 = W.nominal_params(plan, 1)
    pa, pb = perturbed(p0, plan, 5, 0.1), perturbed(p0, plan, 6, 0.1)
    ana.getNLL(torch.tensor(pa, device="cuda"))
    ana.getIntensity(torch.tensor(pb, device="cuda"), 1)
    ref = _flat(O.amplitudes(pa, 1))
    assert (ana.getDataTensor().cpu() - ref).abs().max() <= 1e-10 * ref.abs().max()
    nf = plan.n_free
    v = torch.complex(torch.tensor(pa[:nf]), torch.tensor(pa[nf:2 * nf])).to("cuda")
    ff_v = ana.getFitFractions(v).cpu()
    ff_a = ana.getFitFractions(torch.tensor(pa, device="cuda")).cpu()
    assert torch.allclose(ff_v, ff_a, rtol=1e-12, atol=1e-14)
    ana.getHessian(torch.tensor(pb, device="cuda"))
    ref = _flat(O.amplitudes(pb, 1))
    assert (ana.getDataTensor().cpu() - ref).abs().max() <= 1e-10 * ref.abs().max()
