"""Fit service with continuous batching (ctpwa_amd.serve) on the GPU.

A toy ensemble runs once with all fits in the mix (up to 8 slots, fits admitted as they
arrive and retired as they finish) and once with one slot (every fit alone).  Each fit's
result — minimum, NLL, number of evaluations — must be the same in both runs: the
batched evaluation reproduces each likelihood bit for bit, so the trajectories agree
exactly.  Both minimisers: Newton with the GPU Hessian (every fit must converge, EDM
below the tolerance) and L-BFGS.
"""
import numpy as np
import pytest

from helpers import plan_of
from ctpwa_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", ["newton", "lbfgs"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_service_mix_equals_alone(precision, method):
    from ctpwa_amd.serve import FitService, toy_ensemble
    plan = plan_of("jpsi_3pi")
    s = W.generate_samples("jpsi_3pi", W.device_intensity, n_data=60_000, n_phsp=200_000, n_bkg=0)
    x_nom = W.nominal_params(plan, 1)
    print("case service", precision, method, flush=True)
    kw = dict(n_min=2_000, n_max=20_000, phsp_factor=5, rate=2000.0, x_nominal=x_nom, max_evals=150)
    jobs = toy_ensemble(plan, s["phsp"], s["data"], 10, seed=3, **kw)
    svc = FitService(plan.to_text(), precision=precision, max_slots=8, method=method)
    mix = svc.run(jobs)
    svc.close()
    alone_svc = FitService(plan.to_text(), precision=precision, max_slots=1, method=method)
    alone = alone_svc.run(toy_ensemble(plan, s["phsp"], s["data"], 10, seed=3, **kw))
    alone_svc.close()
    sm = mix.summary()
    print("mix", sm, flush=True)
    print("alone", alone.summary(), flush=True)
    assert sm["fits"] == 10 and alone.summary()["fits"] == 10
    assert sm["mean_batch"] > 1.5                       # fits did share rounds
    assert len(mix.latencies) == sum(r.outcome.n_evals for r in mix.fits)
    a = {r.job_id: r.outcome for r in alone.fits}
    for r in mix.fits:
        o = a[r.job_id]
        assert r.outcome.n_evals == o.n_evals
        assert r.outcome.f == o.f
        assert np.array_equal(r.outcome.x, o.x)
    # the fits move off their start and lower the NLL
    assert all(r.outcome.n_iter > (0 if method == "newton" else 3) for r in mix.fits)
    if method == "newton":
        assert sm["converged"] == 10 and sm["hessians"] > 0
        for r in mix.fits:
            assert r.outcome.edm < 1e-5 and np.all(np.isfinite(r.outcome.errors[2 * plan.n_free:][:6]))
            assert r.outcome.n_hess == a[r.job_id].n_hess
            assert np.array_equal(r.outcome.errors, a[r.job_id].errors, equal_nan=True)
