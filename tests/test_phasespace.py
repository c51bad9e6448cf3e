"""GENBOD phase space and the event text format."""
import numpy as np
import pytest

from ctpwa_amd import phasespace as PS
from ctpwa_amd import workloads as W


def test_genbod_kinematics():
    ev = W.phase_space(20000, seed=4)
    assert ev.shape == (20000, 3, 4)
    tot = ev.sum(axis=1)
    assert np.max(np.abs(tot[:, 1:])) < 1e-12
    assert np.max(np.abs(tot[:, 0] - W.SQRT_S)) < 1e-12
    m = np.sqrt(np.maximum(ev[..., 0] ** 2 - np.sum(ev[..., 1:] ** 2, axis=-1), 0))
    assert np.allclose(m, W.MASSES[None, :], atol=1e-12)


def test_genbod_is_uniform_in_the_dalitz_plane():
    """Uniform three-body phase space is flat in (m12^2, m23^2): compare the m12^2
    projection with the analytic Dalitz-strip length."""
    M, m1, m2, m3 = W.SQRT_S, *W.MASSES
    ev = W.phase_space(200000, seed=9)
    p12 = ev[:, 0] + ev[:, 1]
    s12 = p12[:, 0] ** 2 - np.sum(p12[:, 1:] ** 2, axis=1)
    lo, hi = (m1 + m2) ** 2, (M - m3) ** 2
    edges = np.linspace(lo, hi, 21)
    hist, _ = np.histogram(s12, bins=edges)

    def strip(s):
        E2 = (s - m1 ** 2 + m2 ** 2) / (2 * np.sqrt(s))
        E3 = (M ** 2 - s - m3 ** 2) / (2 * np.sqrt(s))
        p2 = np.sqrt(np.maximum(E2 ** 2 - m2 ** 2, 0))
        p3 = np.sqrt(np.maximum(E3 ** 2 - m3 ** 2, 0))
        return 4 * p2 * p3

    xs = np.linspace(lo, hi, 20001)
    dens = strip(xs)
    cdf = np.concatenate([[0], np.cumsum(0.5 * (dens[1:] + dens[:-1]) * np.diff(xs))])
    expect = np.diff(np.interp(edges, xs, cdf)) / cdf[-1] * len(s12)
    chi2 = np.sum((hist - expect) ** 2 / expect)
    assert chi2 < 60          # 20 bins


def test_seeded_reproducibility():
    a = W.phase_space(1000, seed=3)
    b = W.phase_space(1000, seed=3)
    assert np.array_equal(a, b)


def test_dat_round_trip(tmp_path):
    ev = W.phase_space(500, seed=1)
    f = tmp_path / "e.dat"
    PS.write_dat(str(f), ev)
    back = PS.read_dat(str(f), 3)
    assert np.array_equal(back, ev)
    with pytest.raises(ValueError):
        PS.read_dat(str(f), 7)


def test_accept_reject_follows_weights():
    rng = np.random.default_rng(0)
    x = rng.random(200000)
    cand = x[:, None, None] * np.ones((1, 3, 4))
    sel = PS.accept_reject(cand, x, 10 ** 9, seed=2, i_max=1.0)
    m = sel[:, 0, 0]
    # accepted density ~ 2x on [0, 1]: mean 2/3
    assert abs(m.mean() - 2 / 3) < 0.01
