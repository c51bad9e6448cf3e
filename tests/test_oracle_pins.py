"""Pins of the CPU oracle against references independent of it.

* Clebsch-Gordan coefficients: sympy's CG (Condon-Shortley convention).
* Y_lm: scipy.special.sph_harm_y.
* Wigner small-d: the matrix exponential exp(-i beta J_y) built from the J_+ ladder.
* factorial: math.gamma.
* Blatt-Weisskopf factors and propagators: closed forms at special points.
* node tensor: the spin-0-daughter case computed from scipy Y_lm and sympy CG.
* likelihood: invariances (global phase and scale of the couplings, independence of
  the couplings with one amplitude), central finite differences of the gradient and
  the Hessian, Hessian symmetry, and a numpy brute-force assembly on a tiny sample.
"""
import math

import numpy as np
import pytest
import scipy.linalg as la
from scipy.special import sph_harm_y
from sympy import Rational
from sympy.physics.quantum.cg import CG

from helpers import plan_of, rel, samples


def _half(x2):
    return Rational(x2, 2)


def test_factorial(oracle_lib):
    O = oracle_lib
    for n in range(0, 12):
        assert O.factorial(n) == math.factorial(n)
    for k in range(0, 12):
        n = k + 0.5
        assert O.factorial(n) == pytest.approx(math.gamma(n + 1), rel=1e-14)


@pytest.mark.parametrize("dj1,dj2", [(1, 1), (3, 1), (3, 3), (2, 2), (5, 3), (4, 2), (5, 5), (7, 3)])
def test_clebsch_gordan_vs_sympy(oracle_lib, dj1, dj2):
    O = oracle_lib
    j1x2, j2x2 = dj1 - 1, dj2 - 1
    for Jx2 in range(abs(j1x2 - j2x2), j1x2 + j2x2 + 1, 2):
        for m1 in range(-j1x2, j1x2 + 1, 2):
            for m2 in range(-j2x2, j2x2 + 1, 2):
                M = m1 + m2
                if abs(M) > Jx2:
                    continue
                ref = float(CG(_half(j1x2), _half(m1), _half(j2x2), _half(m2), _half(Jx2), _half(M)).doit())
                got = O.cg(dj1, m1, dj2, m2, Jx2 + 1, M)
                assert got == pytest.approx(ref, abs=1e-13)


def test_clebsch_gordan_selection_rules(oracle_lib):
    O = oracle_lib
    assert O.cg(3, 2, 3, 0, 3, 0) == 0.0        # M != m1 + m2
    assert O.cg(3, 0, 3, 0, 7, 0) == 0.0        # J > j1 + j2


def _stirling(n):
    """The reference's factorial for integers >= 12 (src/Amplitude.cu:64-74)."""
    return math.sqrt(2 * math.pi * n) * (n / math.e) ** n * (1 + 1 / (12 * n) + 1 / (288 * n * n))


@pytest.mark.parametrize("l", range(0, 7))
def test_ylm_vs_scipy(oracle_lib, l):
    """Y_lm matches scipy; where l + |m| >= 12 the reference's normalisation carries its
    Stirling factorial, which the oracle reproduces exactly."""
    rng = np.random.default_rng(l)
    for _ in range(5):
        th, ph = rng.uniform(0, np.pi), rng.uniform(-np.pi, np.pi)
        for m in range(-l, l + 1):
            ref = complex(sph_harm_y(l, m, th, ph))
            if l + abs(m) >= 12:
                ref *= math.sqrt(math.factorial(l + abs(m)) / _stirling(l + abs(m)))
            got = oracle_lib.ylm(l, m, th, ph)
            assert abs(got - ref) < 1e-12


def _jy(j2):
    j = j2 / 2
    ms = np.array([(j2 - 2 * i) / 2 for i in range(j2 + 1)])
    d = len(ms)
    jp = np.zeros((d, d))
    for i in range(1, d):
        m = ms[i]
        jp[i - 1, i] = np.sqrt(j * (j + 1) - m * (m + 1))
    return (jp - jp.T) / 2j


@pytest.mark.parametrize("dim", [2, 3, 4, 5])
def test_wigner_small_d_vs_expm(oracle_lib, dim):
    """The reference's DFunc(m1, m2, beta) is <j m2| exp(-i beta J_y) |j m1>."""
    for beta in (0.3, 1.1, 2.9, -0.7):
        D = la.expm(-1j * beta * _jy(dim - 1))
        for i in range(dim):
            for k in range(dim):
                m1, m2 = dim - 1 - 2 * i, dim - 1 - 2 * k
                assert oracle_lib.dfunc(dim, m1, m2, beta) == pytest.approx(D[k, i].real, abs=1e-12)
                assert abs(D[k, i].imag) < 1e-12


@pytest.mark.parametrize("L", range(0, 7))
def test_blatt_weisskopf_closed_forms(oracle_lib, L):
    O = oracle_lib
    for q0 in (0.1, 0.35, 1.2):
        assert O.bf(L, q0, q0, 3.0) == pytest.approx(1.0, rel=1e-14)
    z, z0 = 0.4 * 3.0, 0.3 * 3.0
    if L == 1:
        assert O.bf(1, 0.4, 0.3, 3.0) == pytest.approx(math.sqrt((1 + z0 ** 2) / (1 + z ** 2)), rel=1e-14)
    if L == 2:
        assert O.bf(2, 0.4, 0.3, 3.0) == pytest.approx(
            math.sqrt((9 + 3 * z0 ** 2 + z0 ** 4) / (9 + 3 * z ** 2 + z ** 4)), rel=1e-14)
    assert O.bf(L, 0.4, 0.3, 0.0) == 1.0


def test_bwr_on_pole(oracle_lib):
    """At m = m0 (q = q0) the width is Gamma0 and BWR = i / (m0 Gamma0)."""
    m0, g0 = 0.775, 0.149
    for L in (0, 1, 2):
        z = oracle_lib.bwr(m0, m0, g0, L, 0.36, 0.36, 3.0)
        assert z.real == pytest.approx(0.0, abs=1e-12)
        assert z.imag == pytest.approx(1.0 / (m0 * g0), rel=1e-13)


def test_bwr_off_pole_formula(oracle_lib):
    m, m0, g0, L, q, q0, d = 0.9, 0.775, 0.149, 1, 0.41, 0.36, 3.0
    bf2 = (1 + (q0 * d) ** 2) / (1 + (q * d) ** 2)
    gam = g0 * (q / q0) ** 3 * (m0 / m) * bf2
    ref = 1.0 / (m0 ** 2 - m ** 2 - 1j * m0 * gam)
    assert abs(oracle_lib.bwr(m, m0, g0, L, q, q0, d) - ref) < 1e-12 * abs(ref)


def _breakup(m, m1, m2):
    return math.sqrt(max((m * m - (m1 + m2) ** 2) * (m * m - (m1 - m2) ** 2), 0.0)) / (2 * m)


def test_gs_on_pole_and_derivative_of_h(oracle_lib):
    """Gounaris-Sakurai (Phys. Rev. Lett. 21 (1968) 244): at m = m0 the dispersive term
    vanishes and F = i (1 + d g0/m0) / (m0 g0); h'(m0) = dh/dm^2 at m0 (checked against a
    central difference of h)."""
    mpi, m0, g0 = 0.13957, 0.775, 0.149
    q0 = _breakup(m0, mpi, mpi)
    d = (3 / math.pi) * mpi ** 2 / q0 ** 2 * math.log((m0 + 2 * q0) / (2 * mpi)) + m0 / (2 * math.pi * q0) \
        - mpi ** 2 * m0 / (math.pi * q0 ** 3)
    got = oracle_lib.gs(m0, m0, g0, q0, q0, mpi)
    ref = 1j * (1 + d * g0 / m0) / (m0 * g0)
    assert abs(got - ref) < 1e-12 * abs(ref)

    def h(m):
        q = _breakup(m, mpi, mpi)
        return (2 / math.pi) * (q / m) * math.log((m + 2 * q) / (2 * mpi))
    eps = 1e-5
    dh_num = (h(math.sqrt(m0 ** 2 + eps)) - h(math.sqrt(m0 ** 2 - eps))) / (2 * eps)
    h0 = h(m0)
    dh_closed = h0 * (1 / (8 * q0 ** 2) - 1 / (2 * m0 ** 2)) + 1 / (2 * math.pi * m0 ** 2)
    assert abs(dh_num - dh_closed) < 1e-8 * abs(dh_closed)


@pytest.mark.parametrize("m", [0.45, 0.7, 0.9, 1.3])
def test_gs_matches_textbook_form(oracle_lib, m):
    """Independent evaluation of the GS propagator: the dispersive term written with the
    numerical derivative of h, Gamma(m) = g0 (q/q0)^3 m0/m."""
    mpi, m0, g0 = 0.13957, 0.775, 0.149
    q0, q = _breakup(m0, mpi, mpi), _breakup(m, mpi, mpi)

    def h(x):
        k = _breakup(x, mpi, mpi)
        return (2 / math.pi) * (k / x) * math.log((x + 2 * k) / (2 * mpi))
    eps = 1e-6
    dh = (h(math.sqrt(m0 ** 2 + eps)) - h(math.sqrt(m0 ** 2 - eps))) / (2 * eps)
    f = g0 * m0 ** 2 / q0 ** 3 * (q ** 2 * (h(m) - h(m0)) + (m0 ** 2 - m ** 2) * q0 ** 2 * dh)
    d = (3 / math.pi) * mpi ** 2 / q0 ** 2 * math.log((m0 + 2 * q0) / (2 * mpi)) + m0 / (2 * math.pi * q0) \
        - mpi ** 2 * m0 / (math.pi * q0 ** 3)
    gam = g0 * (q / q0) ** 3 * m0 / m
    ref = (1 + d * g0 / m0) / (m0 ** 2 - m ** 2 + f - 1j * m0 * gam)
    got = oracle_lib.gs(m, m0, g0, q, q0, mpi)
    assert abs(got - ref) < 2e-7 * abs(ref)


def test_flatte_formula(oracle_lib):
    m, m0 = 1.02, 0.99
    g = np.array([0.2, 0.4])
    ch = np.array([[0.1396, 0.1396], [0.4937, 0.4937]])
    s = m * m
    rho = []
    for ma, mb in ch:
        f1 = 1 - (ma + mb) ** 2 / s
        f2 = 1 - (ma - mb) ** 2 / s
        rho.append(np.sqrt(complex(f1)) * math.sqrt(max(f2, 0.0)))
    ref = 1.0 / (m0 ** 2 - s - 1j * (m0 / 2) * sum(gi * r for gi, r in zip(g, rho)))
    assert abs(oracle_lib.flatte(m, m0, g, ch) - ref) < 1e-12 * abs(ref)


def test_node_tensor_spin0_daughters(oracle_lib):
    """J -> 0 + 0 through L = J: amp[s] = 10 qs^L / sqrt(2J+1) <00;00|00> <00;L s|J s> Y_L^s(theta, -phi)."""
    rng = np.random.default_rng(3)
    m1, m2 = 0.13975, 0.13975
    for J in (1, 2):
        for _ in range(3):
            p = rng.uniform(0.1, 0.6)
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            p1 = np.array([math.hypot(p, m1), *(p * d)])
            p2 = np.array([math.hypot(p, m2), *(-p * d)])
            amp = oracle_lib.node_tensor(p1, 1, p2, 1, 2 * J + 1, 1, J)
            th, ph = math.acos(d[2]), math.atan2(d[1], d[0])
            for si in range(2 * J + 1):
                s = J - si
                cg = float(CG(0, 0, 0, 0, 0, 0).doit())
                cg2 = float(CG(0, 0, J, s, J, s).doit())
                ref = 10 * p ** J / math.sqrt(2 * J + 1) * cg * cg2 * complex(sph_harm_y(J, s, th, -ph))
                assert abs(amp[si, 0, 0] - ref) < 1e-12 * max(1.0, abs(ref))


# ---------------------------------------------------------------------------
# likelihood-level pins
# ---------------------------------------------------------------------------
def _small(oracle_lib, name, n_p=3000, n_d=600, n_b=150, seed=3):
    from ctpwa_amd import workloads as W
    plan = plan_of(name)
    ph, da, bk = samples(n_p, n_d, n_b if name == "jpsi_3pi" else 0, seed)
    wb = None if bk is None else np.full(bk.shape[0], 0.8)
    O = oracle_lib.Oracle(plan.to_text(), 3, ph, da, bk, w_bkg=wb)
    p0 = W.nominal_params(plan, None if name == "jpsi_rho" else 1)
    return plan, O, p0, (ph, da, bk, wb)


def _rotate(p, plan, phi, scale=1.0):
    n = plan.n_free
    v = (p[:n] + 1j * p[n:2 * n]) * scale * np.exp(1j * phi)
    q = p.copy()
    q[:n], q[n:2 * n] = v.real, v.imag
    return q


def test_global_phase_and_scale_invariance(oracle_lib):
    plan, O, p0, _ = _small(oracle_lib, "jpsi_3pi")
    nll0, g0 = O.nll_grad(p0)
    for phi, sc in ((0.7, 1.0), (2.1, 3.5), (-1.3, 0.2)):
        nll, _ = O.nll_grad(_rotate(p0, plan, phi, sc))
        assert nll == pytest.approx(nll0, rel=1e-12, abs=1e-9)
    # d/dphi L(e^{i phi} v) = 0: gradient orthogonal to i v
    n = plan.n_free
    dv = np.concatenate([-p0[n:2 * n], p0[:n]])
    assert abs(np.dot(g0[:2 * n], dv)) < 1e-9 * np.linalg.norm(g0[:2 * n]) * np.linalg.norm(dv)


def test_single_amplitude_coupling_independence(oracle_lib):
    plan, O, p0, _ = _small(oracle_lib, "jpsi_rho")
    assert plan.n_amp == 1
    nll0, g0 = O.nll_grad(p0)
    q = p0.copy()
    q[0], q[1] = 2.7, -1.1
    nll1, _ = O.nll_grad(q)
    assert nll1 == pytest.approx(nll0, rel=1e-12)
    assert np.max(np.abs(g0[:2])) < 1e-9 * np.max(np.abs(g0[2:]))


@pytest.mark.parametrize("name", ["jpsi_rho", "jpsi_3pi"])
def test_gradient_and_hessian_finite_differences(oracle_lib, name):
    plan, O, p0, _ = _small(oracle_lib, name)
    p0 = p0.copy()
    p0[: 2 * plan.n_free] += 0.05
    nll, g = O.nll_grad(p0)
    nll_h, g_h, H = O.hessian(p0)
    assert nll_h == pytest.approx(nll, rel=1e-13)
    assert rel(g_h, g) < 1e-12
    fd = np.zeros_like(p0)
    fdH = np.zeros_like(H)
    for i in range(len(p0)):
        h = 1e-6 * max(1.0, abs(p0[i]))
        pp, pm = p0.copy(), p0.copy()
        pp[i] += h
        pm[i] -= h
        (np_, gp), (nm, gm) = O.nll_grad(pp), O.nll_grad(pm)
        fd[i] = (np_ - nm) / (2 * h)
        fdH[:, i] = (gp - gm) / (2 * h)
    assert rel(g, fd) < 1e-6
    assert rel(H, fdH) < 1e-6
    assert np.max(np.abs(H - H.T)) <= 1e-10 * np.max(np.abs(H))


def test_numpy_brute_force_assembly(oracle_lib):
    """NLL and coupling gradient assembled in numpy from the oracle's amplitudes."""
    plan, O, p0, (ph, da, bk, wb) = _small(oracle_lib, "jpsi_3pi", n_p=200, n_d=50, n_b=20, seed=8)
    n = plan.n_free
    v = p0[:n] + 1j * p0[n:2 * n]
    va = np.array([plan.amp_ratio[a] * v[plan.amp_free[a]] for a in range(plan.n_amp)])

    def I_of(cls):
        A = O.amplitudes(p0, cls)
        S = np.einsum("epa,a->ep", A, va)
        return A, S, np.sum(np.abs(S) ** 2, axis=1)

    Ap, Sp, Ip = I_of(0)
    Ad, Sd, Id = I_of(1)
    Ab, Sb, Ib = I_of(2)
    Neff = len(Id) - 0.8 * len(Ib)
    pf = Ip.mean()
    nll = -np.sum(np.log(Id)) + 0.8 * np.sum(np.log(Ib)) + Neff * np.log(pf)
    gv = (-np.einsum("epa,ep,e->a", Ad.conj(), Sd, 1 / Id) + 0.8 * np.einsum("epa,ep,e->a", Ab.conj(), Sb, 1 / Ib)
          + Neff / (pf * len(Ip)) * np.einsum("epa,ep->a", Ap.conj(), Sp))
    gp = np.zeros(n, dtype=complex)
    for a in range(plan.n_amp):
        gp[plan.amp_free[a]] += plan.amp_ratio[a] * gv[a]
    nll_o, g_o = O.nll_grad(p0)
    assert nll_o == pytest.approx(nll, rel=1e-12)
    assert rel(g_o[:n], 2 * gp.real) < 1e-10
    assert rel(g_o[n:2 * n], 2 * gp.imag) < 1e-10
