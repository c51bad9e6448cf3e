"""Pins of the oracle's spin transforms against explicit Lorentz-boost composition.

The reference carries a daughter's spin states from the mother's rest frame into the
frame the momenta are given in with two closed forms, MassiveTrans and MasslessTrans
(src/Amplitude.cu:331-470), applied in the node tensor (src/Amplitude.cu:570-635).
These tests rebuild the same objects from first principles: 4x4 pure boosts L(p)
taking (m, 0, 0, 0) to p, the Thomas-Wigner rotation R_W = L(p)^-1 L(P) L(p*) (p* the
daughter's momentum in the mother's rest frame), its spin-j representation
exp(-i psi n.J) from the angular-momentum matrices, Clebsch-Gordan coefficients from
sympy and Y_lm from scipy.  Nothing here shares code or formulas with the oracle.

* massive daughters: MassiveTrans(P, p) = D^j(R_W^-1) to rounding, for every spin
  up to 2 and boosts up to gamma ~ 10; identity when P and p are collinear;
* the node tensor with massive spinning daughters, evaluated in a moving frame, is
  10 x the complex conjugate of the canonical two-body amplitude
  q^L / sqrt(2J+1) <j1 s1 j2 s2|S s1+s2> <S s1+s2 L M|J s> Y_L^M(theta*, phi*)
  carried into the moving frame by D^j1(R_W1) x D^j2(R_W2) (the reference evaluates
  Y at -phi and applies the transforms transposed, i.e. the conjugate amplitude);
* massless daughters: D(phi*, theta*, 0) MasslessTrans(P, p) is the m -> 0 limit of
  MassiveTrans(P, p_m) D(phi_F, theta_F, 0) (the little-group phase between the two
  helicity frames), and the node tensor with exactly massless daughters is the limit
  of the massive one in those helicity bases.  For half-integer spin the limit is
  fixed only up to the sign of the SU(2) lift: the reference takes the phase angle
  in (-pi, pi), the continuous limit may take the other branch.
"""
import math

import numpy as np
import pytest
import scipy.linalg as la
from scipy.special import sph_harm_y
from sympy import Rational
from sympy.physics.quantum.cg import CG

METRIC = np.diag([1.0, -1.0, -1.0, -1.0])


def _jmats(dim):
    """J_x, J_y, J_z in the basis m = j, j-1, ..., -j (Condon-Shortley)."""
    j = (dim - 1) / 2
    ms = np.array([j - i for i in range(dim)])
    jp = np.zeros((dim, dim))
    for i in range(1, dim):
        jp[i - 1, i] = math.sqrt(j * (j + 1) - ms[i] * (ms[i] + 1))
    return (jp + jp.T) / 2, (jp - jp.T) / 2j, np.diag(ms)


def _euler(a, b, g, dim):
    """D^j(alpha, beta, gamma) = exp(-i alpha J_z) exp(-i beta J_y) exp(-i gamma J_z)."""
    _, jy, jz = _jmats(dim)
    return la.expm(-1j * a * jz) @ la.expm(-1j * b * jy) @ la.expm(-1j * g * jz)


def _spin_rep(R, dim):
    """exp(-i psi n.J) for the 3x3 rotation R = rotation by psi about n, |psi| < pi."""
    w, V = la.eig(R)
    n = np.real(V[:, np.argmin(np.abs(w - 1.0))])
    n /= la.norm(n)
    c = (np.trace(R) - 1.0) / 2.0
    s = 0.5 * ((R[2, 1] - R[1, 2]) * n[0] + (R[0, 2] - R[2, 0]) * n[1] + (R[1, 0] - R[0, 1]) * n[2])
    psi = math.atan2(s, c)
    jx, jy, jz = _jmats(dim)
    return la.expm(-1j * psi * (n[0] * jx + n[1] * jy + n[2] * jz))


def _mass(p):
    return math.sqrt(max(p @ METRIC @ p, 0.0))


def _four(m, v):
    v = np.asarray(v, dtype=np.float64)
    return np.array([math.sqrt(m * m + v @ v), *v])


def _boost(p):
    """Pure boost taking (m, 0, 0, 0) to p."""
    m = _mass(p)
    g = p[0] / m
    b = p[1:] / p[0]
    B = np.eye(4)
    B[0, 0] = g
    B[0, 1:] = B[1:, 0] = g * b
    bb = b @ b
    if bb > 0:
        B[1:, 1:] += (g - 1.0) * np.outer(b, b) / bb
    return B


def _unboost(p):
    """L(p)^-1 = L of the momentum with the spatial part reversed."""
    return _boost(np.array([p[0], -p[1], -p[2], -p[3]]))


def _wigner(P, ps):
    """R_W = L(p)^-1 L(P) L(p*) (3x3) and p = L(P) p*, for a daughter p* at rest frame of P."""
    p = _boost(P) @ ps
    RW = _unboost(p) @ _boost(P) @ _boost(ps)
    g = max(P[0] / _mass(P), p[0] / _mass(p))
    tol = 1e-9 + 1e-13 * g * g            # rounding of the 4x4 products grows like gamma^2
    assert abs(RW[0, 0] - 1.0) < tol and np.abs(RW[0, 1:]).max() < tol
    return RW[1:, 1:], p


def _angles(v):
    v = v / la.norm(v)
    return math.acos(max(-1.0, min(1.0, v[2]))), math.atan2(v[1], v[0])


def _cg(dj1, m1, dj2, m2, dJ, M):
    h = lambda x: Rational(x, 2)
    return float(CG(h(dj1 - 1), h(m1), h(dj2 - 1), h(m2), h(dJ - 1), h(M)).doit())


def _canonical(P, p1s, m2, dj1, dj2, dj, dS, L):
    """Canonical amplitude with the daughters' spin states defined in the frame of P:
    daughter momenta (p1, p2) in that frame and amp[s][n1][n2]."""
    p2s = _four(m2, -p1s[1:])
    th, ph = _angles(p1s[1:])
    q = la.norm(p1s[1:])
    a0 = np.zeros((dj, dj1, dj2), dtype=complex)
    for si in range(dj):
        s = dj - 1 - 2 * si
        for i1 in range(dj1):
            s1 = dj1 - 1 - 2 * i1
            for i2 in range(dj2):
                s2 = dj2 - 1 - 2 * i2
                M2 = s - s1 - s2
                if abs(M2) > 2 * L:
                    continue
                y = complex(sph_harm_y(L, M2 // 2, th, ph))
                a0[si, i1, i2] = (q ** L / math.sqrt(dj) * _cg(dj1, s1, dj2, s2, dS, s1 + s2)
                                  * _cg(dS, s1 + s2, 2 * L + 1, M2, dj, s) * y)
    R1, p1 = _wigner(P, p1s)
    R2, p2 = _wigner(P, p2s)
    D1, D2 = _spin_rep(R1, dj1), _spin_rep(R2, dj2)
    return p1, p2, np.einsum("no,mp,sop->snm", D1, D2, a0)


def _rand_config(rng, M, m, pmax=2.0, qmax=0.6):
    P = _four(M, rng.normal(size=3) * pmax / math.sqrt(3))
    d = rng.normal(size=3)
    ps = _four(m, d / la.norm(d) * rng.uniform(0.05, qmax))
    return P, ps


@pytest.mark.parametrize("dim", [2, 3, 4, 5])
def test_massive_transform_is_inverse_wigner_rotation(oracle_lib, dim):
    rng = np.random.default_rng(10 + dim)
    for t in range(12):
        # up to gamma ~ 10 for the mother and ~ 4 for the daughter in its rest frame
        P, ps = _rand_config(rng, rng.uniform(0.8, 3.5), rng.uniform(0.1, 1.5), pmax=8.0, qmax=2.0)
        R, p = _wigner(P, ps)
        T = oracle_lib.massive_trans(P, p, dim)
        ref = _spin_rep(R.T, dim)
        assert np.abs(T - ref).max() < 1e-11, (t, np.abs(T - ref).max())
        # a sign error in psi would give D(R_W) instead
        assert np.abs(T - _spin_rep(R, dim)).max() > 1e-6
    # collinear: no rotation
    P = _four(3.0, [0.3, -0.4, 1.2])
    p = _four(0.5, [0.6, -0.8, 2.4])
    assert np.abs(oracle_lib.massive_trans(P, p, dim) - np.eye(dim)).max() < 1e-14


@pytest.mark.parametrize("dim", [2, 3, 4])
def test_massless_transform_is_massless_limit(oracle_lib, dim):
    rng = np.random.default_rng(20 + dim)
    for t in range(10):
        P, _ = _rand_config(rng, 2.5, 0.0)
        d = rng.normal(size=3)
        d /= la.norm(d)
        got = []
        for m in (1e-3, 1e-4):
            ps = _four(m, d * 0.7)
            _, p = _wigner(P, ps)
            th_s, ph_s = _angles(ps[1:])
            th_f, ph_f = _angles(p[1:])
            lim = oracle_lib.massive_trans(P, p, dim) @ _euler(ph_f, th_f, 0.0, dim)
            p0 = np.array([la.norm(p[1:]), *p[1:]])
            ml = _euler(ph_s, th_s, 0.0, dim) @ oracle_lib.massless_trans(P, p0, dim)
            err = np.abs(lim - ml).max()
            if dim % 2 == 0:
                err = min(err, np.abs(lim + ml).max())
            got.append(err)
        # converges linearly in the mass: O(1) mistakes (a sign of the phase, swapped
        # frame axes) stay O(1)
        assert got[0] < 2e-3 and got[1] < 2e-4, (t, got)
        assert abs(np.abs(np.diag(oracle_lib.massless_trans(P, p0, dim))) - 1.0).max() < 1e-14


# (dj1, dj2, dj, dS, L, m1, m2): spin dimensions 2j+1, including half-integer spins
MASSIVE_NODES = [
    (2, 2, 3, 3, 0, 0.938, 0.938),
    (2, 2, 3, 1, 1, 0.938, 0.938),
    (3, 1, 3, 3, 1, 0.775, 0.1396),
    (3, 1, 3, 3, 2, 0.775, 0.1396),
    (3, 3, 1, 1, 0, 0.775, 0.782),
    (2, 3, 2, 2, 0, 0.938, 0.775),
    (2, 3, 4, 4, 2, 0.938, 0.775),
    (3, 3, 5, 5, 2, 0.775, 0.782),
    (4, 2, 3, 3, 1, 1.232, 0.938),
]


@pytest.mark.parametrize("node", MASSIVE_NODES)
def test_node_tensor_massive_daughters_in_moving_frame(oracle_lib, node):
    dj1, dj2, dj, dS, L, m1, m2 = node
    rng = np.random.default_rng(hash(node) % 1000)
    for _ in range(3):
        P, p1s = _rand_config(rng, 3.1, m1, pmax=3.0, qmax=0.5)
        p1, p2, ref = _canonical(P, p1s, m2, dj1, dj2, dj, dS, L)
        amp = oracle_lib.node_tensor(p1, dj1, p2, dj2, dj, dS, L)
        assert np.abs(ref).max() > 1e-3
        assert np.abs(amp - 10.0 * np.conj(ref)).max() < 1e-12 * max(1.0, np.abs(ref).max())


# exactly massless vectors: integer (a, b, c, d) with a^2 + b^2 + c^2 = d^2, scaled by 2^-k
QUADS = [(1, 2, 2, 3), (2, 3, 6, 7), (1, 4, 8, 9), (4, 4, 7, 9), (2, 6, 9, 11), (6, 6, 7, 11), (3, 4, 12, 13)]


def _exact_massless(rng):
    a, b, c, d = QUADS[rng.integers(len(QUADS))]
    v = np.array([a, b, c], dtype=np.float64) * rng.choice([-1.0, 1.0], size=3)
    rng.shuffle(v)
    return np.array([d, *v]) * 2.0 ** -int(rng.integers(2, 5))


MASSLESS_NODES = [
    (3, 1, 3, 3, 1, 0.0, 0.547),     # photon as first daughter
    (3, 1, 3, 3, 2, 0.0, 0.547),
    (1, 3, 3, 3, 1, 0.547, 0.0),     # photon as second daughter
    (3, 3, 3, 3, 0, 0.0, 0.0),       # two photons
    (3, 3, 1, 1, 0, 0.0, 0.0),
    (2, 1, 2, 2, 0, 0.0, 0.938),     # massless spin 1/2
    (1, 2, 2, 2, 1, 0.938, 0.0),
    (2, 3, 4, 4, 2, 0.0, 0.775),
]


@pytest.mark.parametrize("node", MASSLESS_NODES)
def test_node_tensor_massless_daughters_are_the_limit(oracle_lib, node):
    dj1, dj2, dj, dS, L, m1, m2 = node
    rng = np.random.default_rng(hash(node) % 1000 + 7)
    for _ in range(3):
        p1 = _exact_massless(rng) if m1 == 0 else _four(m1, rng.normal(size=3) * 0.5)
        p2 = _exact_massless(rng) if m2 == 0 else _four(m2, rng.normal(size=3) * 0.5)
        amp = oracle_lib.node_tensor(p1, dj1, p2, dj2, dj, dS, L)
        eps = 1e-5
        q1 = _four(eps, p1[1:]) if m1 == 0 else p1
        q2 = _four(eps, p2[1:]) if m2 == 0 else p2
        P = q1 + q2
        p1s = _unboost(P) @ q1
        _, _, ref = _canonical(P, p1s, _mass(q2), dj1, dj2, dj, dS, L)
        lim = 10.0 * np.conj(ref)
        # massless daughters are quoted in their helicity frame in the moving frame
        if m1 == 0:
            th, ph = _angles(p1[1:])
            lim = np.einsum("kn,skm->snm", _euler(ph, th, 0.0, dj1), lim)
        if m2 == 0:
            th, ph = _angles(p2[1:])
            lim = np.einsum("km,snk->snm", _euler(ph, th, 0.0, dj2), lim)
        scale = np.abs(amp).max()
        assert scale > 1e-2
        err = np.abs(amp - lim).max()
        if (dj1 % 2 == 0 and m1 == 0) or (dj2 % 2 == 0 and m2 == 0):
            err = min(err, np.abs(amp + lim).max())
        assert err < 2e-3 * scale, (node, err, scale)
