"""Custom lineshape expressions: parser, and the oracle's evaluation pinned by
equivalences with built-in lineshapes and by finite differences (CPU)."""
import math

import numpy as np
import pytest

from helpers import plan_of, samples, perturbed, rel
from ctpwa_amd import customexpr as ce
from ctpwa_amd import workloads as W
from ctpwa_amd.config import ConfigError

RHO_BWR = "  rho770:  {J: 1, P: -1, model: BWR, parameters: [0.775, 0.149], free: [0, 1]}"


def _custom(expr, names="[p0, p1]"):
    return W.JPSI_3PI.replace(
        RHO_BWR, '  rho770:  {J: 1, P: -1, model: Custom, paramValues: [0.775, 0.149], paramNames: %s, '
                 'free: [0, 1], expression: "%s"}' % (names, expr))


def _ctx():
    return ce.Context(["a", "b"], [1], 3.0, 0.14, 0.14)


@pytest.mark.parametrize("text,expected", [
    ("1 + 2 * 3", ["+", "n:1.0", "*", "n:2.0", "n:3.0"]),
    ("2 ^ 3 ^ 2", ["^", "n:2.0", "^", "n:3.0", "n:2.0"]),          # right associative
    ("-a * b", ["*", "neg", "p:0", "p:1"]),
    ("(1 - a) / b - 1j", ["-", "/", "-", "n:1.0", "p:0", "p:1", "I"]),
    ("exp(i*m) + Re(q)", ["+", "f:exp", "*", "I", "v:m", "f:re", "v:q"]),
    ("L * d", ["*", "n:1.0", "n:3.0"]),
    ("q0", ["q0", "p:0", "n:0.14", "n:0.14"]),
])
def test_parser(text, expected):
    assert ce.serialize(ce.parse(text, _ctx())) == expected


@pytest.mark.parametrize("text", ["a +", "foo(1)", "p3p * 2", "Flatte(a, b, 1)", "Flatte(a, b, c, 1, 2)",
                                  "BWR(a, b, c)", "(a", "a $ b"])
def test_parser_rejects(text):
    with pytest.raises(ConfigError):
        ce.parse(text, _ctx())


def test_L_must_be_unique_at_the_node():
    with pytest.raises(ConfigError):
        ce.parse("L", ce.Context(["a"], [0, 2], 3.0, 0.14, 0.14))


def _oracles(oracle_lib, texts, seed=3):
    ph, da, bk = samples(5000, 1200, 300, seed)
    wb = np.full(bk.shape[0], 0.8)
    out = []
    for t in texts:
        plan = plan_of(t)
        out.append((plan, oracle_lib.Oracle(plan.to_text(), 3, ph, da, bk, w_bkg=wb)))
    return out


def test_custom_bwr_equals_builtin(oracle_lib):
    """BWR(p0, p1, 1, 3.0) * Bf(1, q, q0(p0), 3.0) is the built-in BWR with its vertex
    barrier (rho -> pi pi has L = 1): same NLL, gradient and Hessian."""
    (pc, oc), (pb, ob) = _oracles(oracle_lib, [_custom("BWR(p0, p1, 1, 3.0) * Bf(1, q, q0(p0), 3.0)"),
                                              W.JPSI_3PI])
    p = perturbed(W.nominal_params(pb, 1), pb, 2)
    nc, gc, Hc = oc.hessian(p)
    nb, gb, Hb = ob.hessian(p)
    assert abs(nc - nb) <= 1e-12 * abs(nb)
    assert rel(gc, gb) <= 1e-10
    assert rel(Hc, Hb) <= 1e-9


F0_BWR = "  f0_980:  {J: 0, P: 1,  model: BWR, parameters: [0.99, 0.05],   free: [0, 1]}"


def test_custom_flatte_equals_builtin(oracle_lib):
    """Flatte(m0, g1, g2, ma1, mb1, ma2, mb2) in an expression is the built-in two-channel
    Flatte (f0 -> pi pi has L = 0, so the built-in model's vertex barrier is 1): same NLL,
    gradient and Hessian, the couplings free."""
    chans = "[[0.13957, 0.13957], [0.493677, 0.493677]]"
    built = W.JPSI_3PI.replace(F0_BWR, "  f0_980:  {J: 0, P: 1, model: Flatte, parameters: [0.99, 0.2, 0.4], "
                                       "free: [0, 1, 2], channels: %s}" % chans)
    cust = W.JPSI_3PI.replace(F0_BWR, '  f0_980:  {J: 0, P: 1, model: Custom, paramValues: [0.99, 0.2, 0.4], '
                                      'paramNames: [m0, g1, g2], free: [0, 1, 2], '
                                      'expression: "Flatte(m0, g1, g2, 0.13957, 0.13957, 0.493677, 0.493677)"}')
    (pc, oc), (pb, ob) = _oracles(oracle_lib, [cust, built])
    p = perturbed(W.nominal_params(pb, 1), pb, 6)
    nc, gc, Hc = oc.hessian(p)
    nb, gb, Hb = ob.hessian(p)
    assert abs(nc - nb) <= 1e-12 * abs(nb)
    assert rel(gc, gb) <= 1e-10
    assert rel(Hc, Hb) <= 1e-9


def test_custom_daughter_momentum_variables(oracle_lib):
    """p1p, p1e, p1costheta, p1phi and p2... are the daughters' |p|, E, p_z/|p| and
    atan2(p_y, p_x) in the event frame: a BW times a factor built from them equals the
    built-in BW amplitude times that factor evaluated in numpy on the events."""
    f = "(1 + 0.3*p1costheta - 0.2*p2p + 0.1*p1e*p2e + 0.05*sin(p1phi - p2phi) + 0.02*p2costheta)"
    cust = W.JPSI_RHO.replace("model: BWR, parameters: [0.775, 0.149], free: [0, 1]",
                              'model: Custom, paramValues: [0.775, 0.149], free: [0, 1], '
                              'expression: "p0^2/(p0^2 - m^2 - i*p0*p1) * %s"' % f)
    bw = W.JPSI_RHO.replace("model: BWR", "model: BW")
    pc, pb = plan_of(cust), plan_of(bw)
    ph, da, _ = samples(2000, 500, 0, 17)
    Oc = oracle_lib.Oracle(pc.to_text(), 3, ph, da)
    Ob = oracle_lib.Oracle(pb.to_text(), 3, ph, da)
    p = perturbed(W.nominal_params(pb, None), pb, 2)
    Ac, Ab = Oc.amplitudes(p, 1), Ob.amplitudes(p, 1)
    # rho -> pi+ pi-: daughter 1 = pip (column 0), daughter 2 = pim (column 1)
    def var(ev):
        P = np.linalg.norm(ev[:, 1:], axis=1)
        return P, ev[:, 0], ev[:, 3] / P, np.arctan2(ev[:, 2], ev[:, 1])
    p1p, p1e, p1c, p1f = var(da[:, 0])
    p2p, p2e, p2c, p2f = var(da[:, 1])
    fac = 1 + 0.3 * p1c - 0.2 * p2p + 0.1 * p1e * p2e + 0.05 * np.sin(p1f - p2f) + 0.02 * p2c
    m0 = p[2 * pb.n_free]
    ref = m0 ** 2 * Ab * fac[:, None, None]
    assert np.abs(Ac - ref).max() <= 1e-12 * np.abs(ref).max()


def test_custom_bw_is_scaled_builtin(oracle_lib):
    """p0^2/(p0^2 - m^2 - i*p0*p1) = p0^2 x BW(p0, p1): the rho(770) amplitudes differ by
    the factor p0^2, the others not at all."""
    bw = W.JPSI_3PI.replace(RHO_BWR, "  rho770:  {J: 1, P: -1, model: BW, parameters: [0.775, 0.149], free: [0, 1]}")
    (pc, oc), (pb, ob) = _oracles(oracle_lib, [_custom("p0^2/(p0^2 - m^2 - i*p0*p1)"), bw])
    p = perturbed(W.nominal_params(pb, 1), pb, 5)
    Ac, Ab = oc.amplitudes(p, 1), ob.amplitudes(p, 1)
    m0 = p[2 * pb.n_free + [t.name for t in pb.theta].index("rho770_mass")]
    rho = [a for a, n in enumerate(pb.amp_names) if "rho770" in n]
    other = [a for a in range(pb.n_amp) if a not in rho]
    assert np.allclose(Ac[:, :, rho], m0 ** 2 * Ab[:, :, rho], rtol=1e-12, atol=0)
    assert np.array_equal(Ac[:, :, other], Ab[:, :, other])


def test_custom_functions_finite_differences(oracle_lib):
    """Every operator and function of the language in one expression; the oracle's
    gradient and Hessian against central differences of its NLL and gradient."""
    expr = ("exp(-p1*m) * (cos(p0*m) + i*sin(p1)) / sqrt(p0^2 + m^2) * abs(p0 - i*p1) "
            "+ log(p0) * conj(m + i*p1)^2 - Im(p0 * (q + 1j)) + Re(csqrt(p1 - q0))")
    ((pc, oc),) = _oracles(oracle_lib, [_custom(expr)])
    p = perturbed(W.nominal_params(pc, 1), pc, 7)
    _, g, H = oc.hessian(p)
    j0 = 2 * pc.n_free + [t.name for t in pc.theta].index("rho770_p0")
    for j in (j0, j0 + 1):
        h = 1e-6 * abs(p[j])
        pp, pm = p.copy(), p.copy()
        pp[j] += h
        pm[j] -= h
        fp, gp = oc.nll_grad(pp)
        fm, gm = oc.nll_grad(pm)
        assert abs((fp - fm) / (2 * h) - g[j]) <= 1e-6 * max(abs(g[j]), 1.0)
        assert np.max(np.abs((gp - gm) / (2 * h) - H[j])) <= 1e-5 * np.max(np.abs(H[j]))


def test_fused_kernel_loads_custom_factors():
    """Custom factors are computed by their own generated kernel (ctpwa_custom); the fused
    pass only loads their values and derivatives, so the complex-dual code never shares the
    fused pass's registers."""
    from ctpwa_amd import _native
    src = _native.generate_source(plan_of(_custom("exp(-p1*m) * cos(p0*m) / sqrt(p0^2 + m^2)")).to_text(), "fp64")
    assert "ctpwa_custom(" in src
    body = src[src.index("void eval_body("):src.index('extern "C" __global__')]
    assert "cd_" not in body and "cfb[" in body
    assert "cd_exp(" in src[src.index("ctpwa_custom("):]


def test_custom_modules_compile_for_gfx950():
    """A model with Custom lineshapes compiles two modules with hipRTC: the fused pass with
    its first-order factor kernel, and the Hessian's second-order factor kernel."""
    from ctpwa_amd import _native
    plan = plan_of(_custom("exp(-p1*m) * cos(p0*m) / sqrt(p0^2 + m^2) + BWR(p0, p1, 1, 3.0)"))
    base = plan_of(W.JPSI_3PI)
    assert _native.compile_check(plan.to_text(), "fp64") > _native.compile_check(base.to_text(), "fp64")


# X(1-) -> n nbar: the waves at the X node have L = 0 and L = 2 (S = 1)
PER_WAVE_L = """\
Particles:
  Jpsi: {J: 1, P: -1, mass: 3.0969, polarization: [1, -1]}
  n1:  {J: 1/2, P: 1, mass: 0.93957}
  n2:  {J: 1/2, P: -1, mass: 0.93957}
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
"""
X_BWR = "  X2000: {J: 1, P: -1, model: BWR, parameters: [2.0, 0.3], free: [0, 1]}"
# the built-in BWR: running width with the node's lowest L, vertex barrier with each wave's L
X_CUSTOM = ("  X2000: {J: 1, P: -1, model: Custom, paramValues: [2.0, 0.3], paramNames: [p0, p1], "
            "free: [0, 1], expression: 'BWR(p0, p1, 0, 3.0) * Bf(L, q, q0(p0), 3.0)'}")
PER_WAVE_MASSES = (0.93957, 0.93957, 0.13498)


def test_per_wave_L_splits_the_block_and_equals_builtin(oracle_lib):
    """A Custom expression using L at a node whose waves differ in L (src/CustomExpr.cu:413,
    evaluated per wave at run time in the reference) is parsed once per run of waves with
    one L; the runs are blocks over consecutive SL combinations of one partial wave.  The
    Custom spelling of the built-in BWR then gives the built-in model's NLL, gradient and
    Hessian, amplitude list, partial waves and SL vector counts."""
    from test_identical_cpu import _sample
    from ctpwa_amd import _native
    with pytest.raises(ce.PerWaveL):
        ce.parse("BW(a, 0.1) * Bf(L, q, 0.3, 3.0)", ce.Context(["a"], [0, 2], 3.0, 0.9, 0.9))
    assert X_BWR in PER_WAVE_L
    pb, pc = plan_of(PER_WAVE_L), plan_of(PER_WAVE_L.replace(X_BWR, X_CUSTOM))
    assert len(pc.blocks) == len(pb.blocks) + 1
    xb = [b for b, B in enumerate(pc.blocks) if B.res[0].name == "X2000"]
    assert [(pc.blocks[b].c0, pc.blocks[b].nsl, pc.blocks[b].wave) for b in xb] == [(0, 1, xb[0]), (1, 1, xb[0])]
    L_at = [pc.topologies[pc.blocks[b].topo].sl_combos[pc.blocks[b].c0][1][1] for b in xb]
    assert L_at == [0, 2]
    assert pc.amp_names == pb.amp_names and pc.n_amp == pb.n_amp
    assert len(pc.waves) == len(pb.waves) and pc.wave_of()[xb[1]] == pc.wave_of()[xb[0]]
    assert "bslc0 " in pc.to_text() and "bwave " in pc.to_text()
    counts = [0] * len(pc.waves)
    for w, B in zip(pc.wave_of(), pc.blocks):
        counts[w] += B.nsl
    assert counts == [B.nsl for B in pb.blocks]
    # the native plan reader accepts the split plan and builds its kernels
    assert _native.compile_check(pc.to_text(), "fp64") > 0
    ph, da = _sample(3000, 1, PER_WAVE_MASSES), _sample(800, 2, PER_WAVE_MASSES)
    ob = oracle_lib.Oracle(pb.to_text(), 3, ph, da)
    oc = oracle_lib.Oracle(pc.to_text(), 3, ph, da)
    p = perturbed(W.nominal_params(pb, 1), pb, 2)
    nb, gb, Hb = ob.hessian(p)
    nc, gc, Hc = oc.hessian(p)
    assert abs(nc - nb) <= 1e-12 * abs(nb)
    assert rel(gc, gb) <= 1e-10
    assert rel(Hc, Hb) <= 1e-9
    # a planted mistake: one L for every wave moves the NLL
    bad = plan_of(PER_WAVE_L.replace(X_BWR, X_CUSTOM.replace("Bf(L,", "Bf(0,")))
    nx, _ = oracle_lib.Oracle(bad.to_text(), 3, ph, da).nll_grad(p)
    assert abs(nx - nb) > 1e-6 * abs(nb)


def _b1_case():
    """J/psi -> gamma pi0 pi0 with b1(1235)(1+) -> pi0 gamma (waves with L = 0 and 2) next to
    the omega, the two pi0 identical: the split runs get permuted copies."""
    from test_identical_cpu import BOSONS
    b1 = "  b1_1235:  {J: 1, P: 1, model: BWR, parameters: [1.23, 0.14], free: [0, 1]}"
    tb = BOSONS.replace("        - [J: 1, P: -1]: [omega782]\n",
                        "        - [J: 1, P: -1]: [omega782]\n        - [J: 1, P: 1]: [b1_1235]\n")
    tb = tb.replace("Constraints:", b1 + "\nConstraints:")
    tc = tb.replace(b1, "  b1_1235:  {J: 1, P: 1, model: Custom, paramValues: [1.23, 0.14], paramNames: [p0, p1], "
                        "free: [0, 1], expression: 'BWR(p0, p1, 0, 3.0) * Bf(L, q, q0(p0), 3.0)'}")
    return tb, tc, (0.0, 0.13498, 0.13498)


def test_per_wave_L_with_identical_particles(oracle_lib):
    """The runs of a split block are permuted like any block: each copy keeps its run's
    first SL combination and sums into the same partial wave.  The photon daughter is a
    configured massless final-state particle (mass 0 in q0).  Same NLL, gradient and Hessian
    as the built-in BWR."""
    from test_identical_cpu import _sample
    tb, tc, masses = _b1_case()
    pb, pc = plan_of(tb), plan_of(tc)
    runs = [b for b, B in enumerate(pc.blocks) if B.res[0].name == "b1_1235" and B.sigma == 0]
    assert len(runs) > 1 and len({pc.blocks[b].wave for b in runs}) == 1
    for b, B in enumerate(pc.blocks):
        if B.sigma:
            S = pc.blocks[B.src]
            assert (B.c0, B.nsl, B.wave, B.res[0].expr) == (S.c0, S.nsl, S.wave, S.res[0].expr)
    assert pc.amp_names == pb.amp_names and len(pc.waves) == len(pb.waves)
    ph, da = _sample(3000, 1, masses), _sample(800, 2, masses)
    p = perturbed(W.nominal_params(pb, 1), pb, 2)
    nb, gb, Hb = oracle_lib.Oracle(pb.to_text(), 3, ph, da).hessian(p)
    nc, gc, Hc = oracle_lib.Oracle(pc.to_text(), 3, ph, da).hessian(p)
    assert abs(nc - nb) <= 1e-12 * abs(nb)
    assert rel(gc, gb) <= 1e-10
    assert rel(Hc, Hb) <= 1e-9
