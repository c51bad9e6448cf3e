"""config.yml reader and amplitude enumeration against hand-derived tables.

The (S, L) lists follow the textbook rules |j2-j3| <= S <= j2+j3, |j1-S| <= L <= j1+S,
P1 = P2 P3 (-1)^L (the reference's Amp2BD::ComSL, src/AmpGen.cu:69-129); the expected
tables below are worked out by hand for each decay.
"""
import pytest

from ctpwa_amd.config import ConfigError, ConfigParser
from ctpwa_amd.model import _polarization, build_plan, com_sl
from ctpwa_amd import workloads as W


def _plan(text):
    return build_plan(ConfigParser("", text=text))


@pytest.mark.parametrize("spins,parities,expected", [
    ((3, 1, 3), (-1, -1, -1), [(3, 1)]),                     # J/psi -> pi rho
    ((3, 1, 1), (-1, -1, 1), []),                           # J/psi -> pi f0: parity forbidden
    ((3, 1, 5), (-1, -1, 1), [(5, 2)]),                     # J/psi -> pi f2
    ((3, 1, 1), (-1, -1, -1), [(1, 1)]),                    # rho -> pi pi
    ((3, 3, 1), (1, -1, -1), [(3, 0), (3, 2)]),             # K1(1+) -> K* pi
    ((3, 3, 1), (-1, 1, -1), [(3, 0), (3, 2)]),             # J/psi -> K1 K
    ((3, 3, 1), (-1, -1, -1), [(3, 1)]),                    # J/psi -> K* K
    ((5, 1, 1), (1, -1, -1), [(1, 2)]),                     # f2 -> pi pi
    ((3, 3, 3), (-1, -1, -1), [(1, 1), (3, 1), (5, 1), (5, 3)]),   # 1- -> 1- 1-
])
def test_com_sl(spins, parities, expected):
    assert com_sl(spins, parities) == expected


def test_com_sl_filters():
    allw = com_sl((3, 3, 3), (-1, -1, -1))
    assert com_sl((3, 3, 3), (-1, -1, -1), maxL=1) == [sl for sl in allw if sl[1] <= 1]
    assert com_sl((3, 3, 3), (-1, -1, -1), sl_filter=[(5, 3)]) == [(5, 3)]
    assert com_sl((3, 1, 1), (-1, -1, 1), p_break=True) == [(1, 1)]
    # identical bosons: (-1)^(L+S) = +1
    assert com_sl((1, 3, 3), (1, -1, -1), identical=True, boson=True) == [(1, 0), (5, 2)]


def test_single_chain_plan():
    plan = _plan(W.JPSI_RHO)
    assert (plan.n_amp, plan.n_free, plan.n_theta, plan.npol) == (1, 1, 2, 2)
    assert plan.polmap == [0, 2]
    t = plan.topologies[0]
    assert [n.sl_list for n in t.nodes] == [[(3, 1)], [(1, 1)]]
    assert [n.lmin for n in t.nodes] == [1, 1]
    assert plan.param_names[1:] == ["rho770_mass", "rho770_width"]
    assert [(th.init, th.lower, th.upper) for th in plan.theta] == [
        (0.775, 0.775 - 0.3875, 0.775 + 0.3875), (0.149, 0.149 - 0.0745, 0.149 + 0.0745)]
    assert plan.amp_names == ["Jpsi→pi0+rho770_SL(3,1)_rho770→pip+pim_SL(1,1)"]


def test_three_chain_plan():
    plan = _plan(W.JPSI_3PI)
    assert (plan.n_amp, plan.n_free, plan.n_theta, plan.npol) == (10, 10, 10, 2)
    chains = [t.chain for t in plan.topologies]
    assert chains == ["c1_R_pm", "c1_R_pm", "c1_R_pm", "c1_R_m0", "c1_R_p0"]
    # blocks: rho0 x3, f0 (no wave), f2, rho- x3, rho+ x3
    nsl = [b.nsl for b in plan.blocks]
    assert nsl == [1, 1, 1, 0, 1, 1, 1, 1, 1, 1, 1]
    assert [b.site for b in plan.blocks] == [0, 1, 2, 3, 3, 4, 5, 6, 7, 8, 9]
    # theta slots registered in block order, shared by name across chains
    assert [t.name for t in plan.theta] == [
        "rho770_mass", "rho770_width", "rho1450_mass", "rho1450_width", "rho1700_mass",
        "rho1700_width", "f0_980_mass", "f0_980_width", "f2_1270_mass", "f2_1270_width"]
    assert plan.blocks[5].res[0].slots == plan.blocks[0].res[0].slots == [0, 1]
    assert plan.blocks[8].res[0].slots == [0, 1]
    f2 = plan.topologies[2]
    assert [n.sl_list for n in f2.nodes] == [[(5, 2)], [(1, 2)]]
    assert plan.amp_free == list(range(10))


def test_polarization_map_two_polarised_particles():
    from ctpwa_amd.config import Particle
    parts = [Particle("A", 3, -1, 3.0, polarization_2m=[2, -2]), Particle("b", 2, 1, 0.9, polarization_2m=[1]),
             Particle("c", 1, -1, 0.1)]
    npol, pm = _polarization(parts)
    # dims [3, 2, 1]; keep A in {0, 2}, b in {0}; strides [2, 1, 1]
    assert npol == 2
    assert pm == [0, 4]


def test_compact_multi_mode_and_expanded_forms():
    text = W._PARTICLES + """
Data: {order: [pip, pim, pi0]}
DecayChains:
  dc:
    decay:
      - Jpsi: [pi0, R]
      - R: [pip, pim]
    R:
      - [J: 1, P: -1]: [rho770]
Resonances:
  rho770: {J: 1, P: -1, model: BWR, parameters: [0.775, 0.149]}
"""
    plan = _plan(text)
    assert plan.n_amp == 1 and plan.n_theta == 0
    assert plan.topologies[0].chain == "dc"


def test_unsupported_features_raise():
    with pytest.raises(ConfigError):
        _plan(W.JPSI_RHO.replace("model: BWR", "model: Custom"))


TRANS = W.JPSI_3PI + "Constraints:\n  trans:\n    - [R_m0, R_p0]: -1\n"


def test_trans_folds_chains_onto_their_partner():
    """[R_m0, R_p0]: -1 folds each rho(+) chain onto the rho(-) chain of the same
    resonance with ratio -1 (src/Parameters.cu:586-650): 7 free couplings remain."""
    plan = _plan(TRANS)
    assert plan.n_amp == 10 and plan.n_free == 7
    keys = [b.chain_key for b in plan.blocks if b.nsl]
    # chain key: expanded chain name + "_" + intermediate + "[resonance]" (src/Info.cu:243-247)
    assert keys[4:] == ["c1_R_m0_R_m0[rho770]", "c1_R_m0_R_m0[rho1450]", "c1_R_m0_R_m0[rho1700]",
                        "c1_R_p0_R_p0[rho770]", "c1_R_p0_R_p0[rho1450]", "c1_R_p0_R_p0[rho1700]"]
    assert plan.amp_free == [0, 1, 2, 3, 4, 5, 6, 4, 5, 6]
    assert plan.amp_ratio == [1.0] * 7 + [-1.0] * 3
    assert plan.param_names[4].startswith("Jpsi→pip+rho770")
    assert len(plan.param_names) == 7 + 10


def test_trans_with_split_chains():
    """When a chain holds several amplitudes (J/psi -> K1 K: 2 SL combos x 2) every owner
    amplitude gets its own coupling and the folded chain maps by position."""
    text = W.JPSI_KKPIPI + "Constraints:\n  trans:\n    - [K1p, K1m]: [[0.5, 0.0]]\n"
    plan = _plan(text)
    base = _plan(W.JPSI_KKPIPI)
    assert base.n_free == 20 and plan.n_free == 16
    a_p = [a for a, n in enumerate(plan.amp_names) if "→Km+K1_1270" in n]
    a_m = [a for a, n in enumerate(plan.amp_names) if "→Kp+K1_1270" in n]
    assert len(a_p) == len(a_m) == 4
    assert [plan.amp_free[a] for a in a_m] == [plan.amp_free[a] for a in a_p]
    assert all(plan.amp_ratio[a] == 0.5 for a in a_m)
    assert all(plan.amp_ratio[a] == 1.0 for a in a_p)
    assert sorted(set(plan.amp_free)) == list(range(16))


def test_var_equal_merges_slots():
    text = W.JPSI_3PI + "Constraints:\n  var_equal:\n    - [rho1450_width, rho1700_width]\n"
    plan = _plan(text)
    base = _plan(W.JPSI_3PI)
    assert plan.n_theta == base.n_theta - 1
    names = [t.name for t in plan.theta]
    assert "rho1700_width" not in names
    w = names.index("rho1450_width")
    for b in plan.blocks:
        for r in b.res:
            if r.name in ("rho1450", "rho1700"):
                assert r.slots[1] == w
            if r.name == "f2_1270":
                assert r.slots == [names.index("f2_1270_mass"), names.index("f2_1270_width")]


def test_barrier_scope_resolution():
    cfg = ConfigParser("", text=W.JPSI_RHO + "Constraints:\n  bf_d: 5.0\n")
    steps = cfg.decay_chains[0].decay_steps
    assert [s.bf_d for s in steps] == [5.0, 5.0]
    cfg = ConfigParser("", text=W.JPSI_RHO.replace("free: [0, 1]}", "free: [0, 1], bf_d: 4.0, has_bf: false}"))
    steps = cfg.decay_chains[0].decay_steps
    assert steps[0].bf_d == 3.0 and steps[0].has_bf
    assert steps[1].bf_d == 4.0 and not steps[1].has_bf


def test_plan_text_round_trip_fields():
    plan = _plan(W.JPSI_3PI)
    txt = plan.to_text()
    assert txt.startswith("CTPWA_PLAN 1\n") and txt.rstrip().endswith("end")
    assert "namp 10 10" in txt and "ntheta 10" in txt
