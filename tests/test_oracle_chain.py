"""Pins of the oracle's chain contraction and polarisation pick.

The reference multiplies the node tensors of a decay chain in node order, contracting
each node's mother label with the running tensor and appending the node's two daughter
labels (src/AmpGen.cu:1104-1200), then reads the result at the polarisation map's
tensor indices (src/AmpGen.cu:1202-1217; map from initializePolarization,
src/Analysis.cu:4249-4322).  The single-node tensors are pinned elsewhere; these tests
pin their composition against objects built without the oracle:

* J/psi(m = +-1) -> rho0 pi0, rho0 -> pi+ pi- (both L = 1): the contracted tensor is a
  constant times e_m^* . (p_pi- x p_pi+) in the J/psi rest frame (two vectors coupled to
  a vector are their cross product), with |constant| = 100 / (4 pi sqrt 2);
* a three-node cascade J/psi -> K- K1+, K1+ -> K*0 pi+, K*0 -> K+ pi- (the K* spin
  Wigner-rotated inside the moving K1): every SL combination equals the einsum of three
  canonical node amplitudes built from sympy Clebsch-Gordan coefficients, scipy Y_lm and
  explicit boost composition;
* a chain with spin-1/2 final-state particles whose contraction label order differs
  from the particle list: the polarisation index runs over the masked J/psi states and
  the fermion helicities in the contraction's label order, by a hand-written map.
"""
import math

import numpy as np
import pytest

from ctpwa_amd import workloads as W
from helpers import plan_of
from test_oracle_wigner import _canonical, _four

M_PIC, M_PI0, M_K, M_P = 0.13975, 0.13495, 0.493677, 0.938272


def _oracle(oracle_lib, text, events):
    from oracle.oracle import Oracle
    plan = plan_of(text)
    O = Oracle(plan.to_text(), events.shape[1], events, events)
    return plan, O


def _sph(m):
    """e_m of the spherical basis: e_+1 = -(x + i y)/sqrt2, e_0 = z, e_-1 = (x - i y)/sqrt2."""
    s = 1.0 / math.sqrt(2.0)
    return {1: np.array([-s, -1j * s, 0.0]), 0: np.array([0.0, 0.0, 1.0]), -1: np.array([s, -1j * s, 0.0])}[m]


def test_rho_pi_chain_is_the_cross_product(oracle_lib):
    ev = W.phase_space(40, seed=11)
    plan, O = _oracle(oracle_lib, W.JPSI_RHO, ev)
    assert plan.polmap == [0, 2] and len(plan.topologies) == 1
    S = O.slamp(1, 0, len(plan.topologies[0].sl_combos))[0]        # [e, p]
    n = np.cross(ev[:, 1, 1:], ev[:, 0, 1:])                        # p_pi- x p_pi+
    ratio = []
    for p, m in enumerate((1, -1)):
        proj = n @ np.conj(_sph(m))
        ratio.append(S[:, p] / proj)
    ratio = np.concatenate(ratio)
    k = ratio[0]
    assert np.abs(ratio - k).max() < 1e-12 * abs(k)
    assert abs(abs(k) - 100.0 / (4.0 * math.pi * math.sqrt(2.0))) < 1e-12 * abs(k)
    # the pick is not the other order or the m = 0 state
    wrong = S[:, 0] / (n @ np.conj(_sph(-1)))
    assert np.abs(wrong - wrong[0]).max() > 1e-3 * abs(wrong[0])


K1_CHAIN = """\
Particles:
  Jpsi: {J: 1, P: -1, mass: 3.0969, polarization: [1, -1]}
  Kp:  {J: 0, P: -1, mass: 0.493677}
  Km:  {J: 0, P: -1, mass: 0.493677}
  pip: {J: 0, P: -1, mass: 0.13957}
  pim: {J: 0, P: -1, mass: 0.13957}
Data:
  order: [Kp, Km, pip, pim]
DecayChains:
  k1p:
    Jpsi: [[Km, K1p]]
    K1p: [Kst0, pip]
    Kst0: [Kp, pim]
    intermediates:
      K1p:
        - [J: 1, P: 1]: [K1_1270]
      Kst0:
        - [J: 1, P: -1]: [Kst892]
Resonances:
  K1_1270:  {J: 1, P: 1, model: BWR, parameters: [1.253, 0.090], free: [0, 1]}
  Kst892:   {J: 1, P: -1, model: BWR, parameters: [0.8955, 0.0473], free: [0, 1]}
"""


def _direction(rng):
    d = rng.normal(size=3)
    return d / np.linalg.norm(d)


def _two_body_q(M, m1, m2):
    return math.sqrt((M * M - (m1 + m2) ** 2) * (M * M - (m1 - m2) ** 2)) / (2.0 * M)


def _k1_events(n, seed, slc):
    """Events of the K1 cascade built node by node, with the three canonical node
    tensors (10 x conj, the reference's convention) of SL combination slc."""
    rng = np.random.default_rng(seed)
    mpi = 0.13957
    evs, tens = [], []
    PJ = np.array([3.0969, 0.0, 0.0, 0.0])
    for _ in range(n):
        mkst = rng.uniform(M_K + mpi + 0.05, 1.4)
        mk1 = rng.uniform(mkst + mpi + 0.05, 3.0969 - M_K - 0.05)
        (s0, l0), (s1, l1), (s2, l2) = slc
        qj = _two_body_q(3.0969, M_K, mk1)
        pkm_s = _four(M_K, _direction(rng) * qj)
        pkm, pk1, t0 = _canonical(PJ, pkm_s, mk1, 1, 3, 3, s0, l0)
        q1 = _two_body_q(mk1, mkst, mpi)
        pkst, ppip, t1 = _canonical(pk1, _four(mkst, _direction(rng) * q1), mpi, 3, 1, 3, s1, l1)
        q2 = _two_body_q(mkst, M_K, mpi)
        pkp, ppim, t2 = _canonical(pkst, _four(M_K, _direction(rng) * q2), mpi, 1, 1, 3, s2, l2)
        evs.append(np.stack([pkp, pkm, ppip, ppim]))
        tens.append([10.0 * np.conj(t0), 10.0 * np.conj(t1), 10.0 * np.conj(t2)])
    return np.array(evs), tens


def test_three_node_cascade_is_the_einsum_of_canonical_nodes(oracle_lib):
    plan = plan_of(K1_CHAIN)
    topo = plan.topologies[0]
    # J/psi -> K- K1 (S=1, L=0,2), K1 -> K* pi (S=1, L=0,2), K* -> K pi (S=0, L=1)
    assert sorted(topo.sl_combos) == sorted(((3, l0), (3, l1), (1, 1)) for l0 in (0, 2) for l1 in (0, 2))
    for c, slc in enumerate(topo.sl_combos):
        ev, tens = _k1_events(12, 100 + c, slc)
        from oracle.oracle import Oracle
        O = Oracle(plan.to_text(), 4, ev, ev)
        S = O.slamp(1, 0, len(topo.sl_combos))[c]
        # labels: [J, K-, K1] -> [J, K-, K*, pi+] -> [J, K-, pi+, K+, pi-]
        exp = np.array([np.einsum("axb,bcy,czw->axyzw", *t).reshape(-1)[[0, 2]] for t in tens])
        scale = np.abs(exp).max()
        assert scale > 1e-3
        assert np.abs(S - exp).max() < 1e-11 * scale, (slc, np.abs(S - exp).max(), scale)
        # the comparison separates a transposed contraction (the K1 tensor's mother and
        # K* axes exchanged)
        transposed = np.array([np.einsum("axb,cby,czw->axyzw", *t).reshape(-1)[[0, 2]] for t in tens])
        assert np.abs(transposed - exp).max() > 1e-6 * scale


SPIN_HALF = """\
Particles:
  Jpsi: {J: 1, P: -1, mass: 3.0969, polarization: [1, -1]}
  p:    {J: 1/2, P: 1, mass: 0.938272}
  pbar: {J: 1/2, P: -1, mass: 0.938272}
  pi0:  {J: 0, P: -1, mass: 0.13495}
Data:
  order: [p, pbar, pi0]
DecayChains:
  nstar:
    Jpsi: [[pbar, Nst]]
    Nst: [p, pi0]
    intermediates:
      Nst:
        - [J: 1/2, P: 1]: [N1440]
Resonances:
  N1440: {J: 1/2, P: 1, model: BWR, parameters: [1.44, 0.35], free: [0, 1]}
"""


def test_final_state_spins_follow_the_label_order(oracle_lib):
    plan = plan_of(SPIN_HALF)
    # particles [Jpsi (3 states, m = +-1 kept), p (2), pbar (2), pi0 (1)]: C-order strides
    # (4, 2, 1, 1), masked index (m, a, b) -> 4 * {0, 2}[m] + 2 a + b
    assert plan.npol == 8
    assert plan.polmap == [0, 1, 2, 3, 8, 9, 10, 11]
    topo = plan.topologies[0]
    rng = np.random.default_rng(7)
    PJ = np.array([3.0969, 0.0, 0.0, 0.0])
    for c, ((s0, l0), (s1, l1)) in enumerate(topo.sl_combos):
        evs, exp = [], []
        for _ in range(10):
            mn = rng.uniform(M_P + M_PI0 + 0.05, 3.0969 - M_P - 0.05)
            ppb, pn, t0 = _canonical(PJ, _four(M_P, _direction(rng) * _two_body_q(3.0969, M_P, mn)), mn, 2, 2, 3, s0, l0)
            pp, ppi, t1 = _canonical(pn, _four(M_P, _direction(rng) * _two_body_q(mn, M_P, M_PI0)), M_PI0, 2, 1, 2, s1, l1)
            evs.append(np.stack([pp, ppb, ppi]))
            # labels [J, pbar, N] -> [J, pbar, p, pi0]
            full = np.einsum("axb,byz->axyz", 10.0 * np.conj(t0), 10.0 * np.conj(t1)).reshape(-1)
            exp.append(full[plan.polmap])
        ev = np.array(evs)
        exp = np.array(exp)
        from oracle.oracle import Oracle
        O = Oracle(plan.to_text(), 3, ev, ev)
        S = O.slamp(1, 0, len(topo.sl_combos))[c]
        scale = np.abs(exp).max()
        assert scale > 1e-3
        assert np.abs(S - exp).max() < 1e-11 * scale, ((s0, l0, s1, l1), np.abs(S - exp).max(), scale)
        # the fermion labels in particle-list order would be a different pick
        swapped = np.array([np.swapaxes(np.asarray(e).reshape(2, 2, 2), 1, 2).reshape(-1) for e in exp])
        assert np.abs(swapped - exp).max() > 1e-6 * scale
