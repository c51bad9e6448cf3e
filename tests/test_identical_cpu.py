"""Identical-particle symmetrisation (Constraints.identical, src/AmpGen.cu:194-403,
1440-1460, src/Info.cu:150-162) on the CPU.

The plan builder expands every chain with identical particles into permuted copies of
its topologies whose amplitudes carry the original couplings times sgn(sigma), so that
the coherent sum holds A(x) + sum_sigma sgn(sigma) A(x_sigma).  Pinned by:

* the selection rule (-1)^(L+S) = +1 (bosons) / -1 (fermions) for identical daughters
  against hand-derived (S, L) tables;
* each permuted copy's amplitudes equal the unexpanded plan's amplitudes evaluated on
  events whose identical particles' momenta were exchanged by hand;
* the symmetrised intensity equals a numpy brute-force sum over the permutation group of
  the unexpanded amplitudes with the fermion signs, and is invariant under every
  permutation of the identical particles' momenta (bosons and fermions, two and three
  members);
* the reference's combination of two groups in one chain (the first group's
  permutations, repeated per non-identity element of the later groups).
"""
import dataclasses
import itertools

import numpy as np
import pytest

from helpers import plan_of
from ctpwa_amd import workloads as W
from ctpwa_amd.model import _identical_permutations, com_sl

BOSONS = """\
Particles:
  Jpsi: {J: 1, P: -1, mass: 3.0969, polarization: [1, -1]}
  gam:  {J: 1, P: -1, mass: 0.0}
  pi01: {J: 0, P: -1, mass: 0.13498}
  pi02: {J: 0, P: -1, mass: 0.13498}
Data:
  order: [gam, pi01, pi02]
  data: [npy, data.npy]
  phsp: [npy, phsp.npy]
DecayChains:
  f:
    Jpsi: [[gam, R_00]]
    R_00: [pi01, pi02]
    intermediates:
      R_00:
        - [J: 0, P: 1]: [f0_980]
        - [J: 2, P: 1]: [f2_1270]
  w:
    Jpsi: [[pi01, R_w]]
    R_w: [pi02, gam]
    intermediates:
      R_w:
        - [J: 1, P: -1]: [omega782]
Resonances:
  f0_980:   {J: 0, P: 1, model: BWR, parameters: [0.99, 0.07], free: [0, 1]}
  f2_1270:  {J: 2, P: 1, model: BWR, parameters: [1.275, 0.185], free: [0, 1]}
  omega782: {J: 1, P: -1, model: BWR, parameters: [0.7827, 0.0085], free: [0, 1]}
Constraints:
  identical: [[pi01, pi02]]
"""

FERMIONS = """\
Particles:
  Jpsi: {J: 1, P: -1, mass: 3.0969, polarization: [1, -1]}
  n1:  {J: 1/2, P: 1, mass: 0.93957}
  n2:  {J: 1/2, P: 1, mass: 0.93957}
  pi0: {J: 0, P: -1, mass: 0.13498}
Data:
  order: [n1, n2, pi0]
  data: [npy, data.npy]
  phsp: [npy, phsp.npy]
DecayChains:
  nstar:
    Jpsi: [[n1, Nst]]
    Nst: [n2, pi0]
    intermediates:
      Nst:
        - [J: 1/2, P: -1]: [N1535]
        - [J: 3/2, P: -1]: [N1520]
  x:
    Jpsi: [[pi0, R_nn]]
    R_nn: [n1, n2]
    intermediates:
      R_nn:
        - [J: 1, P: -1]: [X2000]
Resonances:
  N1535: {J: 1/2, P: -1, model: BWR, parameters: [1.53, 0.15], free: [0, 1]}
  N1520: {J: 3/2, P: -1, model: BWR, parameters: [1.515, 0.11], free: [0, 1]}
  X2000: {J: 1, P: -1, model: BWR, parameters: [2.0, 0.3], free: [0, 1]}
Constraints:
  identical: [[n1, n2]]
"""

THREE = """\
Particles:
  Jpsi: {J: 1, P: -1, mass: 3.0969, polarization: [1, -1]}
  pi01: {J: 0, P: -1, mass: 0.13498}
  pi02: {J: 0, P: -1, mass: 0.13498}
  pi03: {J: 0, P: -1, mass: 0.13498}
Data:
  order: [pi01, pi02, pi03]
  data: [npy, data.npy]
  phsp: [npy, phsp.npy]
DecayChains:
  f:
    Jpsi: [[pi01, R_00]]
    R_00: [pi02, pi03]
    intermediates:
      R_00:
        - [J: 2, P: 1]: [f2_1270]
        - [J: 2, P: 1]: [f2_1525]
Resonances:
  f2_1270: {J: 2, P: 1, model: BWR, parameters: [1.275, 0.185], free: [0, 1]}
  f2_1525: {J: 2, P: 1, model: BWR, parameters: [1.517, 0.086], free: [0, 1]}
Constraints:
  identical: [[pi01, pi02, pi03]]
"""

# (config, final-state masses, indices of the identical final-state particles, fermions)
CASES = {
    "bosons": (BOSONS, (0.0, 0.13498, 0.13498), (1, 2), False),
    "fermions": (FERMIONS, (0.93957, 0.93957, 0.13498), (0, 1), True),
    "three": (THREE, (0.13498, 0.13498, 0.13498), (0, 1, 2), False),
}


def _sample(n, seed, masses):
    ev = W.phase_space(n, seed=seed, masses=masses)
    if masses[0] == 0.0:
        # an exactly massless photon: E slightly below |p| gives mass 0 in every evaluation
        ev[:, 0, 0] = np.sqrt((ev[:, 0, 1:] ** 2).sum(-1)) * (1.0 - 1e-12)
    return ev


def _strip(plan):
    """The plan without its permuted copies (the unsymmetrised amplitudes)."""
    orig = [b for b, blk in enumerate(plan.blocks) if blk.sigma == 0]
    n_topo = 1 + max(plan.blocks[b].topo for b in orig)
    amps = plan.ref_amplitudes()
    assert amps == list(range(len(amps)))
    blocks = [dataclasses.replace(plan.blocks[b], wave=k) for k, b in enumerate(orig)]
    return dataclasses.replace(plan, topologies=plan.topologies[:n_topo], blocks=blocks, n_amp=len(amps),
                               amp_names=plan.amp_names[:len(amps)], amp_free=plan.amp_free[:len(amps)],
                               amp_ratio=plan.amp_ratio[:len(amps)])


def _permuted(ev, idx, perm):
    """Events with the momenta of the identical particles idx permuted: slot idx[i]
    receives the momentum of idx[perm[i]]."""
    out = ev.copy()
    for i, j in enumerate(perm):
        out[:, idx[i]] = ev[:, idx[j]]
    return out


def _coupled(plan, params):
    n = plan.n_free
    p = params[:n] + 1j * params[n:2 * n]
    return np.array([plan.amp_ratio[a] * p[plan.amp_free[a]] for a in range(plan.n_amp)])


def test_selection_rule_for_identical_daughters():
    # f2 -> pi0 pi0: L even only (bosons, S = 0)
    assert com_sl((5, 1, 1), (1, -1, -1), identical=True, boson=True) == [(1, 2)]
    # rho -> pi0 pi0 is forbidden
    assert com_sl((3, 1, 1), (-1, -1, -1), identical=True, boson=True) == []
    assert com_sl((3, 1, 1), (-1, -1, -1)) == [(1, 1)]
    # 1- -> n n (P = +1 each): parity needs L odd; fermions need L + S odd -> S = 0, L = 1
    assert com_sl((3, 2, 2), (-1, 1, 1), identical=True, boson=False) == [(1, 1)]
    assert com_sl((3, 2, 2), (-1, 1, 1)) == [(1, 1), (3, 1)]
    # 0+ -> gamma gamma with identical photons (bosons, S in {0, 1, 2}): L + S even
    assert com_sl((1, 3, 3), (1, -1, -1), identical=True, boson=True) == [(1, 0), (5, 2)]


def test_plan_structure():
    plan = plan_of(BOSONS)
    # chain f (two resonance sets) and chain w: one permuted copy each, sign +1
    assert len(plan.blocks) == 6 and plan.waves == [0, 1, 2] and plan.wave_of() == [0, 1, 2, 0, 1, 2]
    assert plan.n_amp == 2 * len(plan.ref_amplitudes()) and plan.n_free == 8
    assert plan.amp_ratio == [1.0] * plan.n_amp
    assert "bwave 0 1 2 0 1 2" in plan.to_text()
    fplan = plan_of(FERMIONS)
    n = len(fplan.ref_amplitudes())
    assert fplan.amp_ratio == [1.0] * n + [-1.0] * n
    assert fplan.amp_free[n:] == fplan.amp_free[:n]
    # the copy of chain x reads n2 where the chain names n1 and vice versa
    x = fplan.topologies[fplan.blocks[2].topo].nodes[1]
    xc = fplan.topologies[fplan.blocks[5].topo].nodes[1]
    assert (x.d1_ref, x.d2_ref) == (0, 1) and (xc.d1_ref, xc.d2_ref) == (1, 0)
    tplan = plan_of(THREE)
    assert len(tplan.blocks) == 2 * 6 and tplan.waves == [0, 1]


@pytest.mark.parametrize("name", sorted(CASES))
def test_copies_are_the_amplitudes_of_permuted_events(oracle_lib, name):
    text, masses, idx, _ = CASES[name]
    plan = plan_of(text)
    base = _strip(plan)
    ev = _sample(400, 11, masses)
    params = W.nominal_params(plan, 3)
    A = oracle_lib.Oracle(plan.to_text(), 3, ev, ev[:10]).amplitudes(params, 0)
    ident = tuple(range(len(idx)))
    perms = list(itertools.permutations(ident))
    Ab = {p: oracle_lib.Oracle(base.to_text(), 3, _permuted(ev, idx, p), ev[:10]).amplitudes(params, 0)
          for p in perms}
    for w in plan.waves:
        orig = plan.blocks[w]
        sl = slice(orig.site, orig.site + orig.nsl)

        def same(X, Y):
            return np.abs(X - Y).max() <= 1e-12 * np.abs(Y).max()

        # permutations that give the same base amplitudes (a sibling swap of identical
        # daughters leaves them unchanged) form one class
        classes = []
        for p in perms:
            for c in classes:
                if same(Ab[p][:, :, sl], Ab[c[0]][:, :, sl]):
                    c.append(p)
                    break
            else:
                classes.append([p])
        want = {frozenset(c): len([p for p in c if p != ident]) for c in classes}
        got = {frozenset(c): 0 for c in classes}
        for blk in plan.blocks:
            if blk.sigma == 0 or blk.wave != w:
                continue
            X = A[:, :, blk.site:blk.site + blk.nsl]
            hit = [c for c in classes if same(X, Ab[c[0]][:, :, sl])]
            assert len(hit) == 1, (w, blk.sigma)
            got[frozenset(hit[0])] += 1
        # every non-identity permutation of the group has exactly one copy
        assert got == want, (w, got, want)


@pytest.mark.parametrize("name", sorted(CASES))
def test_symmetrised_intensity(oracle_lib, name):
    text, masses, idx, fermions = CASES[name]
    plan = plan_of(text)
    base = _strip(plan)
    ev = _sample(300, 12, masses)
    params = W.nominal_params(plan, 5)
    O = oracle_lib.Oracle(plan.to_text(), 3, ev, ev[:10])
    I = O.intensity(params, 0)
    # brute force: sum over the group of sgn(sigma) A(x_sigma) with the base plan
    v = _coupled(base, params)
    tot = 0.0
    for p in itertools.permutations(range(len(idx))):
        inv = sum(1 for i in range(len(p)) for j in range(i + 1, len(p)) if p[i] > p[j])
        sg = -1.0 if (fermions and inv % 2) else 1.0
        Ob = oracle_lib.Oracle(base.to_text(), 3, _permuted(ev, idx, p), ev[:10])
        tot = tot + sg * np.einsum("epa,a->ep", Ob.amplitudes(params, 0), v)
    ref = np.sum(np.abs(tot) ** 2, axis=1)
    assert np.abs(I - ref).max() <= 1e-11 * np.abs(ref).max()
    # Bose / Fermi symmetry: the intensity does not change under any exchange
    for p in itertools.permutations(range(len(idx))):
        Op = oracle_lib.Oracle(plan.to_text(), 3, _permuted(ev, idx, p), ev[:10])
        Ip = Op.intensity(params, 0)
        assert np.abs(Ip - I).max() <= 1e-10 * np.abs(I).max(), p
    # and the unsymmetrised model is not symmetric (so the check has teeth)
    Ob = oracle_lib.Oracle(base.to_text(), 3, ev, ev[:10])
    Obs = oracle_lib.Oracle(base.to_text(), 3, _permuted(ev, idx, (1, 0) + tuple(range(2, len(idx)))), ev[:10])
    assert np.abs(Ob.intensity(params, 0) - Obs.intensity(params, 0)).max() > 1e-3 * np.abs(I).max()


def test_two_groups_combine_as_the_reference_does():
    """Two groups in one chain: the reference lays each earlier group's complete map over
    the later group's (src/AmpGen.cu:366-385), so the copies are the first group's
    permutations, once per non-identity element of the second group, signs multiplied."""
    from ctpwa_amd.config import DecayStep
    steps = [DecayStep("Jpsi", ["R1", "R2"]), DecayStep("R1", ["a1", "b1"]), DecayStep("R2", ["a2", "b2"]),
             DecayStep("X", ["b3", "c"])]
    final = ["a1", "b1", "a2", "b2", "b3", "c"]
    # group 1: the subtrees R1, R2 (fermion sign); group 2: b1, b2, b3 (bosons)
    out = _identical_permutations(steps, "Jpsi", [(["R1", "R2"], False), (["b1", "b2", "b3"], True)], final)
    assert len(out) == 1 * 5
    for read, sg in out:
        # R1 <-> R2 with their daughters; the b's stay where group 1 put them
        assert read == [2, 3, 0, 1, 4, 5] and sg == -1.0
    # one group alone: all of S_3 but the identity, bosons
    out = _identical_permutations(steps, "Jpsi", [(["b1", "b2", "b3"], True)], final)
    assert sorted(tuple(r[1:5:2] + r[4:5]) for r, _ in out) == sorted(
        p for p in itertools.permutations((1, 3, 4)) if p != (1, 3, 4))
