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


def test_service_isolates_failing_jobs():
    """An oversized job is rejected by admission, a job with a NaN starting point is
    rejected, a job whose events cannot be loaded fails, and a job whose likelihood is not
    finite at its (finite) start fails after its first evaluation; every other fit runs to
    the same result as in a run without them."""
    from ctpwa_amd.serve import FitJob, FitService, toy_ensemble
    plan = plan_of("jpsi_3pi")
    s = W.generate_samples("jpsi_3pi", W.device_intensity, n_data=60_000, n_phsp=300_000, n_bkg=0)
    x_nom = W.nominal_params(plan, 1)
    kw = dict(n_min=2_000, n_max=8_000, phsp_factor=5, rate=2000.0, x_nominal=x_nom, max_evals=60)
    print("case service isolation", flush=True)
    good = toy_ensemble(plan, s["phsp"], s["data"], 6, seed=4, **kw)
    ref_svc = FitService(plan.to_text(), precision="fp64", max_slots=8, method="newton")
    budget = ref_svc.footprint((250_000, 50_000, 0)) - 1
    ref_svc.close()
    t_mid = good[2].arrival
    lo, hi = good[0].lower, good[0].upper
    x_nan = good[0].x0.copy()
    x_nan[3] = np.nan
    # one data event with NaN momenta: the loss is NaN, reported as 1e30 (the reference's
    # rule, src/Analysis.cu:1145-1158), at the starting point
    nan_data = s["data"][:3000].copy()
    nan_data[17, 0, :] = np.nan

    def broken():
        raise OSError("sample file missing")

    small = {"data": s["data"][:3000], "phsp": s["phsp"][:15000]}
    bad = [FitJob(100, {"data": s["data"][:50_000], "phsp": s["phsp"][:250_000]}, good[1].x0, lo, hi, arrival=t_mid),
           FitJob(101, small, x_nan, lo, hi, arrival=t_mid),
           FitJob(102, broken, good[1].x0, lo, hi, arrival=t_mid),
           FitJob(103, {"data": nan_data, "phsp": s["phsp"][:15000]}, good[1].x0, lo, hi, arrival=t_mid)]
    runs = []
    for jobs in (good + bad, toy_ensemble(plan, s["phsp"], s["data"], 6, seed=4, **kw)):
        svc = FitService(plan.to_text(), precision="fp64", max_slots=8, method="newton", hbm_budget=budget)
        runs.append(svc.run(jobs))
        svc.close()
    mixed, clean = runs
    by = {r.job_id: r for r in mixed.fits}
    assert by[100].status == "rejected" and "HBM" in by[100].error
    assert by[101].status == "rejected" and "not finite" in by[101].error
    assert by[102].status == "failed" and "sample file missing" in by[102].error
    assert by[103].status == "failed" and "not finite" in by[103].error, by[103].error
    sm = mixed.summary()
    assert sm["fits"] == 6 and sm["rejected"] == 2 and sm["failed"] == 2
    ref = {r.job_id: r.outcome for r in clean.fits}
    assert len(ref) == 6 and all(r.status == "ok" for r in clean.fits)
    for j, o in ref.items():
        got = by[j]
        assert got.status == "ok"
        assert got.outcome.f == o.f and got.outcome.n_evals == o.n_evals
        assert np.array_equal(got.outcome.x, o.x)
