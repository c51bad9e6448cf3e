"""Table lineshapes (Hist, Interp: src/Resonance.cu:76-188, include/ResModel.cuh:67-85,
153-157, 226-290) on the CPU.

The oracle's lookups are pinned against what interpolation fixes independently:
numpy.interp (an independent piecewise-linear implementation) on random tables inside
the range; exact reproduction of linear data, including the reference's rules outside
the range (Hist: end values, the last bin held at the next-to-last value; Interp on an
even grid: one segment of linear extrapolation below x0 from truncation toward zero,
then the end values; on an uneven grid: linear extrapolation past both ends).  The
configuration reader is checked for the table layouts it writes and for its errors.
"""
import math
import os

import numpy as np
import pytest

from helpers import plan_of, samples
from ctpwa_amd import workloads as W
from ctpwa_amd.config import ConfigError


def _hist_table(lo, hi, vals):
    return [lo, hi, float(len(vals))] + list(vals)


def test_hist_lookup(oracle_lib):
    rng = np.random.default_rng(1)
    lo, hi, n = 0.3, 2.9, 30
    v = rng.uniform(0.0, 2.0, n)
    t = _hist_table(lo, hi, v)
    # inside the first n - 1 bins: linear interpolation of the values at x = n (m - lo) / (hi - lo)
    for m in rng.uniform(lo, lo + (hi - lo) * (n - 2) / n, 200):
        x = n * (m - lo) / (hi - lo)
        assert oracle_lib.hist(t, m) == pytest.approx(np.interp(x, np.arange(n), v), rel=1e-14, abs=1e-15)
    # the last bin holds the next-to-last value; the ends hold the end values
    for m in rng.uniform(lo + (hi - lo) * (n - 1) / n, hi, 20):
        assert oracle_lib.hist(t, m) == v[n - 2]
    assert oracle_lib.hist(t, lo) == v[0] and oracle_lib.hist(t, 0.1) == v[0]
    assert oracle_lib.hist(t, hi) == v[-1] and oracle_lib.hist(t, 5.0) == v[-1]
    assert oracle_lib.hist(t, -0.5) == 0.0 and oracle_lib.hist(t, float("nan")) == 0.0
    # linear data are reproduced exactly between bins
    lin = 0.5 + 0.25 * np.arange(n)
    tl = _hist_table(lo, hi, lin)
    for m in np.linspace(lo + 1e-9, lo + (hi - lo) * (n - 2) / n, 50):
        assert oracle_lib.hist(tl, m) == pytest.approx(0.5 + 0.25 * n * (m - lo) / (hi - lo), rel=1e-13)


def test_interp_lookup_even_grid(oracle_lib):
    rng = np.random.default_rng(2)
    N, x0, dx = 40, 0.25, 0.07
    xs = x0 + dx * np.arange(N)
    re, im = rng.normal(size=N), rng.normal(size=N)
    t = [1.0, float(N), x0, dx] + list(re) + list(im)
    for x in rng.uniform(xs[0], xs[-1], 300):
        z = oracle_lib.interp(t, x)
        assert z.real == pytest.approx(np.interp(x, xs, re), rel=1e-12, abs=1e-12)
        assert z.imag == pytest.approx(np.interp(x, xs, im), rel=1e-12, abs=1e-12)
    # linear data a + b x: reproduced inside, extrapolated less than one step below x0
    # (truncation toward zero keeps segment 0), held at the ends beyond that
    a, b, c, d = 0.3, -1.2, 2.0, 0.5
    tl = [1.0, float(N), x0, dx] + list(a + b * xs) + list(c + d * xs)
    for x in np.concatenate([rng.uniform(xs[0], xs[-1], 50), x0 - dx * rng.uniform(0.01, 0.99, 20)]):
        z = oracle_lib.interp(tl, x)
        assert z == pytest.approx(complex(a + b * x, c + d * x), rel=1e-12, abs=1e-12)
    for x in (x0 - 1.5 * dx, x0 - 10.0):
        assert oracle_lib.interp(tl, x) == pytest.approx(complex(a + b * xs[0], c + d * xs[0]), rel=1e-13)
    for x in (xs[-1], xs[-1] + 0.3 * dx, xs[-1] + 5.0):
        assert oracle_lib.interp(tl, x) == pytest.approx(complex(a + b * xs[-1], c + d * xs[-1]), rel=1e-12)


def test_interp_lookup_uneven_grid(oracle_lib):
    rng = np.random.default_rng(3)
    N = 25
    xs = np.sort(rng.uniform(0.2, 3.0, N))
    re, im = rng.normal(size=N), rng.normal(size=N)
    t = [-2.0, float(N)] + list(xs) + list(re) + list(im)
    for x in rng.uniform(xs[0], xs[-1], 300):
        z = oracle_lib.interp(t, x)
        assert z.real == pytest.approx(np.interp(x, xs, re), rel=1e-12, abs=1e-12)
        assert z.imag == pytest.approx(np.interp(x, xs, im), rel=1e-12, abs=1e-12)
    # past the ends the end segments extrapolate linearly: exact for linear data
    a, b, c, d = -0.7, 0.9, 1.5, -2.5
    tl = [-2.0, float(N)] + list(xs) + list(a + b * xs) + list(c + d * xs)
    for x in (xs[0] - 0.4, xs[0] - 3.0, xs[-1] + 0.2, xs[-1] + 4.0, 0.5 * (xs[3] + xs[4])):
        assert oracle_lib.interp(tl, x) == pytest.approx(complex(a + b * x, c + d * x), rel=1e-11, abs=1e-11)


def _write(d, name, text):
    p = os.path.join(d, name)
    with open(p, "w") as f:
        f.write(text)
    return p


def _cfg(resonance_line):
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
    Jpsi: [[pi0, R_pm]]
    R_pm: [pip, pim]
    intermediates:
      R_pm:
        - [J: 1, P: -1]: [X]
Resonances:
  X: %s
""" % resonance_line


def test_config_builds_tables(tmp_path):
    d = str(tmp_path)
    _write(d, "h.txt", "1 2 3\n4 5.5\n")
    plan = plan_of(_cfg("{J: 1, P: -1, model: Hist, file: %s/h.txt, bins: 5, range: [0.3, 1.23456789]}" % d))
    r = plan.blocks[0].res[0]
    # the range passes through "%f" as in the reference (six decimals)
    assert r.model == "Hist" and r.table == [0.3, 1.234568, 5.0, 1.0, 2.0, 3.0, 4.0, 5.5]
    assert "rtab 8 " in plan.to_text() and plan.n_theta == 0
    _write(d, "even.txt", "0.5 1 -1\n0.75 2 -2\n1.0 3 -3\n")
    r = plan_of(_cfg("{J: 1, P: -1, model: Interp, file: %s/even.txt}" % d)).blocks[0].res[0]
    assert r.table == [1.0, 3.0, 0.5, 0.25, 1.0, 2.0, 3.0, -1.0, -2.0, -3.0]
    _write(d, "uneven.txt", "0.5 1 -1\n0.7 2 -2\n1.0 3 -3\n")
    r = plan_of(_cfg("{J: 1, P: -1, model: Interp, file: %s/uneven.txt}" % d)).blocks[0].res[0]
    assert r.table == [-2.0, 3.0, 0.5, 0.7, 1.0, 1.0, 2.0, 3.0, -1.0, -2.0, -3.0]
    # a relative path resolves against the config file's directory when it is not found
    cfgp = _write(d, "c.yml", _cfg("{J: 1, P: -1, model: Interp, file: even.txt}"))
    from ctpwa_amd.config import ConfigParser
    from ctpwa_amd.model import build_plan
    assert build_plan(ConfigParser(cfgp)).blocks[0].res[0].table[0] == 1.0


def test_config_table_errors(tmp_path):
    d = str(tmp_path)
    _write(d, "h.txt", "1 2 3\n")
    with pytest.raises(ConfigError, match="expected 4"):
        plan_of(_cfg("{J: 1, P: -1, model: Hist, file: %s/h.txt, bins: 4, range: [0.3, 1.2]}" % d))
    with pytest.raises(ConfigError, match="bins >= 2"):
        plan_of(_cfg("{J: 1, P: -1, model: Hist, file: %s/h.txt, bins: 1, range: [0.3, 1.2]}" % d))
    with pytest.raises(ConfigError, match="hi > lo"):
        plan_of(_cfg("{J: 1, P: -1, model: Hist, file: %s/h.txt, bins: 3, range: [1.2, 0.3]}" % d))
    with pytest.raises(ConfigError, match="requires file"):
        plan_of(_cfg("{J: 1, P: -1, model: Hist, bins: 3, range: [0.3, 1.2]}"))
    with pytest.raises(ConfigError, match="cannot open"):
        plan_of(_cfg("{J: 1, P: -1, model: Interp, file: %s/missing.txt}" % d))
    _write(d, "one.txt", "0.5 1 2\n")
    with pytest.raises(ConfigError, match="two rows"):
        plan_of(_cfg("{J: 1, P: -1, model: Interp, file: %s/one.txt}" % d))


def test_table_lineshape_likelihood_is_the_tabulated_function(oracle_lib, tmp_path):
    """An Interp table holding a BW propagator on a fine grid gives the BW amplitudes to
    interpolation accuracy up to one constant: the parent's q0 takes the daughter mass 1
    for a resonance without a mass parameter instead of m0 (src/AmpGen.cu:1504-1527),
    which rescales its barrier normalisation; the NLL does not see that constant.  A Hist
    of a constant is the ONE model without its barrier, times the constant."""
    d = str(tmp_path)
    m0, g0 = 0.775, 0.149
    x = np.linspace(0.25, 3.0, 20001)
    bw = 1.0 / (m0 ** 2 - x ** 2 - 1j * m0 * g0)
    np.savetxt(os.path.join(d, "bw.txt"), np.stack([x, bw.real, bw.imag], 1), fmt="%.17g")
    ph, da, _ = samples(2000, 400, 0, 9)
    ref = plan_of(_cfg("{J: 1, P: -1, model: BW, parameters: [%r, %r]}" % (m0, g0)))
    tab = plan_of(_cfg("{J: 1, P: -1, model: Interp, file: %s/bw.txt}" % d))
    p = W.nominal_params(ref, None)
    O1 = oracle_lib.Oracle(ref.to_text(), 3, ph, da)
    O2 = oracle_lib.Oracle(tab.to_text(), 3, ph, da)
    A1, A2 = O1.amplitudes(p, 1), O2.amplitudes(W.nominal_params(tab, None), 1)
    # the top-node barrier with q0 from daughter masses (m0 vs 1) -> a constant ratio
    def q0(m, m1, m2):
        return math.sqrt((m * m - (m1 + m2) ** 2) * (m * m - (m1 - m2) ** 2)) / (2 * m)
    z1, z2 = 3.0 * q0(3.0969, 0.13498, m0), 3.0 * q0(3.0969, 0.13498, 1.0)
    c = math.sqrt((1 + z2 * z2) / (1 + z1 * z1))       # L = 1 at the top node
    assert np.abs(A2 - c * A1).max() <= 1e-5 * np.abs(A1).max()
    assert O2.nll_grad(W.nominal_params(tab, None))[0] == pytest.approx(O1.nll_grad(p)[0], rel=1e-6)
    np.savetxt(os.path.join(d, "flat.txt"), np.full(10, 2.0))
    one = plan_of(_cfg("{J: 1, P: -1, model: ONE, parameters: [0.775], has_bf: false}"))
    # the optional Hist mass parameter enters the parent's q0 as ONE's mass does
    hist = plan_of(_cfg("{J: 1, P: -1, model: Hist, parameters: [0.775], file: %s/flat.txt, bins: 10, "
                        "range: [0.2, 3.0]}" % d))
    A3 = oracle_lib.Oracle(one.to_text(), 3, ph, da).amplitudes(W.nominal_params(one, None), 1)
    A4 = oracle_lib.Oracle(hist.to_text(), 3, ph, da).amplitudes(W.nominal_params(hist, None), 1)
    assert np.abs(A4 - 2.0 * A3).max() <= 1e-12 * np.abs(A3).max()
