"""Amplitude enumeration: config -> model plan.

The plan is the host-side description of the likelihood: decay topologies with their
(S, L) partial waves, resonance blocks, the amplitude order, the free resonance
parameter slots (theta) and the coupling layout.  It follows the reference's
enumeration order exactly, because the amplitude order and the parameter layout are
part of the drop-in contract:

* (S, L) rules: ``Amp2BD::ComSL``                 src/AmpGen.cu:69-129
* SL combinations (last step fastest):            src/AmpGen.cu:489-508
* width convention Lmin per node:                 src/AmpGen.cu:443-454
* chains x J^P combos x resonance combos x SL:    src/Info.cu:43-357,
                                                  src/Analysis.cu:4481-4830
* theta slot registration and sharing by name:    src/AmpGen.cu:2009-2242
* polarisation map (config-particle C order):     src/Analysis.cu:4249-4322
* vspace coupling layout (fit mode 1):            src/Analysis.cu:4382-4478,
                                                  src/Parameters.cu:586-731

The plan is handed to the native library (and to the CPU oracle) as a small
line-oriented text document, see :meth:`Plan.to_text`.
"""
from __future__ import annotations

import copy
import itertools
import os
import warnings
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

from . import customexpr
from .config import ConfigError, ConfigParser, Particle, ResonanceConfig
from .customexpr import PerWaveL

# lineshape model codes shared with the native library
MODEL_CODES = {"BWR": 0, "BW": 1, "ONE": 2, "Flatte": 3, "GS": 4, "Custom": 5, "Hist": 6, "Interp": 7}
# ordered parameter names per model (src/Resonance.cu:14-75)
_MODEL_PARAM_SPECS = {
    "BWR": [("mass", False), ("width", False), ("r", True)],
    "BW": [("mass", False), ("width", False)],
    "ONE": [("mass", False)],
    "Flatte": [("mass", False)],
    "GS": [("mass", False), ("width", False)],
    "Custom": [],
    "Hist": [("mass", True)],
    "Interp": [],
}
_EXTRA_PREFIX = {"Flatte": "g"}


def canonical_model(name: str) -> str:
    for k in MODEL_CODES:
        if k.lower() == name.lower():
            return k
    raise ConfigError("lineshape model '%s' is not supported by this build" % name)


def ordered_params(model: str, values: List[float]) -> Tuple[List[float], List[str]]:
    """Resonance::setParamsFromModel (src/Resonance.cu:355-395)."""
    specs = _MODEL_PARAM_SPECS[model]
    required = sum(1 for _, opt in specs if not opt)
    if len(values) < required:
        raise ConfigError("model '%s' requires at least %d parameter(s), got %d"
                          % (model, required, len(values)))
    names, vals = [], []
    for i, (nm, _) in enumerate(specs):
        if i < len(values):
            names.append(nm)
            vals.append(float(values[i]))
    pref = _EXTRA_PREFIX.get(model, "")
    if pref:
        for i in range(len(specs), len(values)):
            names.append("%s%d" % (pref, i))
            vals.append(float(values[i]))
    return vals, names


def com_sl(spins, parities, p_break=False, identical=False, boson=True, maxL=-1,
           sl_filter=()) -> List[Tuple[int, int]]:
    """Allowed (2S+1, L) for mother -> d1 d2, in the reference's order (src/AmpGen.cu:69-129)."""
    tj1, tj2, tj3 = spins[0] - 1, spins[1] - 1, spins[2] - 1
    p1, p2, p3 = parities
    out = []
    for tS in range(abs(tj2 - tj3), tj2 + tj3 + 1, 2):
        for tL in range(abs(tj1 - tS), tj1 + tS + 1, 2):
            L = tL // 2
            sign = 1 if L % 2 == 0 else -1
            if not (p_break or p1 == p2 * p3 * sign):
                continue
            if identical:
                S = tS // 2
                if (1 if (L + S) % 2 == 0 else -1) != (1 if boson else -1):
                    continue
            if maxL > 0 and L > maxL:
                continue
            if sl_filter and (tS + 1, L) not in set(sl_filter):
                continue
            out.append((tS + 1, L))
    return out


@dataclass
class Node:
    mother: str
    d1: str
    d2: str
    spins: Tuple[int, int, int]
    parities: Tuple[int, int, int]
    sl_list: List[Tuple[int, int]]
    has_bf: bool
    bf_d: float
    masses: Tuple[float, float, float]   # config mass or -1 (intermediates)
    d1_ref: int = 0                      # >=0 final index; -(k+1): mother of node k
    d2_ref: int = 0
    lmin: int = 0


@dataclass
class Topology:
    chain: str
    comb: List[Tuple[str, int, int]]     # (intermediate, 2J+1, P)
    nodes: List[Node]
    sl_combos: List[Tuple[Tuple[int, int], ...]]


@dataclass
class ResInst:
    name: str
    tag: str                 # intermediate name it occupies
    model: str
    params: List[float]
    param_names: List[str]
    channels: List[Tuple[float, float]]
    node: int                # node of the topology whose mother is ``tag``
    slots: List[int] = field(default_factory=list)   # theta slot per param or -1
    expr: List[str] = field(default_factory=list)    # Custom: serialized expression
    table: List[float] = field(default_factory=list)  # Hist / Interp: the shape table


@dataclass
class Block:
    topo: int
    site: int
    nsl: int
    res: List[ResInst]
    chain_key: str
    wave: int = -1           # partial wave: its first block (the block itself unless the wave
                             # is split into runs of one L, or the block is a permuted copy)
    sigma: int = 0           # identical-particle permutation (0: the identity)
    c0: int = 0              # first of the topology's SL combinations that the block's waves use
    src: int = -1            # a permuted copy: the block it copies


@dataclass
class ThetaSlot:
    name: str
    res_name: str
    param_idx: int
    init: float
    lower: float
    upper: float


@dataclass
class Plan:
    particles: List[Particle]
    final_order: List[str]
    initial: str
    npol: int
    polmap: List[int]
    topologies: List[Topology]
    blocks: List[Block]
    n_amp: int
    amp_names: List[str]
    theta: List[ThetaSlot]
    amp_free: List[int]          # amplitude -> free coupling index
    amp_ratio: List[float]       # amplitude -> real ratio (trans folding)
    n_free: int
    param_names: List[str]
    gauss: List[Tuple[int, float, float]] = field(default_factory=list)  # (slot, mu, sigma)

    @property
    def n_theta(self) -> int:
        return len(self.theta)

    @property
    def waves(self) -> List[int]:
        """Blocks that are partial waves (the identity permutation), in block order."""
        return [b for b, blk in enumerate(self.blocks) if blk.wave == b]

    def ref_amplitudes(self) -> List[int]:
        """Amplitudes of the identity permutation (the reference's amplitude list)."""
        return [a for blk in self.blocks if blk.sigma == 0 for a in range(blk.site, blk.site + blk.nsl)]

    def wave_of(self) -> List[int]:
        """Partial-wave position (index into ``waves``) of every block."""
        pos = {b: i for i, b in enumerate(self.waves)}
        return [pos[blk.wave] for blk in self.blocks]

    def n_params(self) -> int:
        return 2 * self.n_free + self.n_theta

    def to_text(self) -> str:
        """Serialise for the native library and the oracle (one record per line)."""
        out = ["CTPWA_PLAN 1"]
        out.append("nfinal %d" % len(self.final_order))
        out.append("npol %d %s" % (self.npol, " ".join(str(x) for x in self.polmap)))
        out.append("ntopo %d" % len(self.topologies))
        for t, topo in enumerate(self.topologies):
            out.append("topo %d %d %d" % (t, len(topo.nodes), len(topo.sl_combos)))
            for k, nd in enumerate(topo.nodes):
                out.append("node %d %d %d %d %d %d %d %r %d %r %r %r" % (
                    k, nd.spins[0], nd.spins[1], nd.spins[2], nd.d1_ref, nd.d2_ref,
                    1 if nd.has_bf else 0, float(nd.bf_d), nd.lmin,
                    float(nd.masses[0]), float(nd.masses[1]), float(nd.masses[2])))
            for c in topo.sl_combos:
                out.append("slc " + " ".join("%d %d" % sl for sl in c))
        out.append("nblock %d" % len(self.blocks))
        for b, blk in enumerate(self.blocks):
            out.append("block %d %d %d %d %d" % (b, blk.topo, blk.site, blk.nsl, len(blk.res)))
            for r in blk.res:
                toks = ["res", str(r.node), str(MODEL_CODES[r.model]), str(len(r.params))]
                toks += [repr(float(x)) for x in r.params]
                toks += [str(s) for s in r.slots]
                toks.append(str(len(r.channels)))
                for ch in r.channels:
                    toks += [repr(float(ch[0])), repr(float(ch[1]))]
                out.append(" ".join(toks))
                if r.model == "Custom":
                    out.append("cexpr %d %s" % (len(r.expr), " ".join(r.expr)))
                if r.table:
                    out.append("rtab %d %s" % (len(r.table), " ".join(repr(float(x)) for x in r.table)))
        out.append("ntheta %d" % len(self.theta))
        for j, th in enumerate(self.theta):
            out.append("theta %d %r %r %r" % (j, th.init, th.lower, th.upper))
        out.append("namp %d %d" % (self.n_amp, self.n_free))
        out.append("amap " + " ".join(str(x) for x in self.amp_free))
        out.append("aratio " + " ".join(repr(float(x)) for x in self.amp_ratio))
        if any(blk.wave != b for b, blk in enumerate(self.blocks)):
            out.append("bwave " + " ".join(str(x) for x in self.wave_of()))
        if any(blk.c0 for blk in self.blocks):
            out.append("bslc0 " + " ".join(str(blk.c0) for blk in self.blocks))
        out.append("ngauss %d" % len(self.gauss))
        for (s, mu, sig) in self.gauss:
            out.append("gauss %d %r %r" % (s, mu, sig))
        out.append("end")
        return "\n".join(out) + "\n"


def _polarization(particles: List[Particle]) -> Tuple[int, List[int]]:
    """initializePolarization (src/Analysis.cu:4249-4322).  Returns (nPol, map)."""
    n = len(particles)
    dims = [p.spin for p in particles]
    keep = []
    any_masked = False
    for p in particles:
        if p.is_polarized():
            any_masked = True
            keep.append([(p.spin - 1 - tm) // 2 for tm in p.polarization_2m])
        else:
            keep.append([])
    orig = [1] * n
    for i in range(n - 2, -1, -1):
        orig[i] = orig[i + 1] * dims[i + 1]
    eff = [len(k) if k else d for k, d in zip(keep, dims)]
    masked = [1] * n
    for i in range(n - 2, -1, -1):
        masked[i] = masked[i + 1] * eff[i + 1]
    npol = 1
    for d in eff:
        npol *= d
    if not any_masked:
        return npol, list(range(npol))
    pm = []
    for out in range(npol):
        rem, idx = out, 0
        for p in range(n):
            j = rem // masked[p]
            rem %= masked[p]
            i = keep[p][j] if keep[p] else j
            idx += i * orig[p]
        pm.append(idx)
    return npol, pm


def _trans_folds(cfg: ConfigParser, chain_keys: List[str]) -> Dict[int, Tuple[int, float]]:
    """``Constraints.trans`` pairs [A, B]: ratio (src/Parameters.cu:586-650,
    src/Info.cu:322-335): every chain key containing B is folded onto the chain key
    containing A that carries the same resonance (the text from the last '['), with the
    real part of the first value as the coupling ratio (1 if none)."""
    folds: Dict[int, Tuple[int, float]] = {}

    def res_name(k: str) -> str:
        i = k.rfind("[")
        return k[i:] if i >= 0 else k

    for c in cfg.constraints:
        if c.type != "trans" or len(c.names) < 2:
            continue
        ratio = c.values[0].real if c.values else 1.0
        a_name, b_name = c.names[0], c.names[1]
        ca: Dict[str, int] = {}
        cb: Dict[str, int] = {}
        for ci, k in enumerate(chain_keys):
            if a_name in k:
                ca[res_name(k)] = ci
            if b_name in k:
                cb[res_name(k)] = ci
        for res in sorted(cb):                           # std::map order
            if res in ca:
                folds[cb[res]] = (ca[res], ratio)
    return folds


def _vspace_map(cfg: ConfigParser, amp_chain_key: List[str], chain_keys: List[str]):
    """Amplitude -> free coupling index and ratio in fit mode 1 (src/Analysis.cu:4382-4478)."""
    n_amp = len(amp_chain_key)
    folds = _trans_folds(cfg, chain_keys)
    active = [ci for ci in range(len(chain_keys)) if ci not in folds]
    amp_chain = [0] * n_amp
    amp_ratio = [1.0] * n_amp
    for a, ck in enumerate(amp_chain_key):
        oc = chain_keys.index(ck)
        fc, r = folds.get(oc, (oc, 1.0))
        amp_chain[a] = active.index(fc)
        amp_ratio[a] = r
    owners: Dict[int, List[int]] = {}
    for a in range(n_amp):
        if abs(amp_ratio[a] - 1.0) < 1e-10:
            owners.setdefault(amp_chain[a], []).append(a)
    if not any(len(v) > 1 for v in owners.values()):
        return amp_chain, amp_ratio, len(active)
    # one free coupling per owner amplitude; folded chains map by position onto the
    # free couplings of the chain they fold into
    orig: Dict[str, List[int]] = {}
    for a, ck in enumerate(amp_chain_key):
        orig.setdefault(ck, []).append(a)
    folded = {ck for ck, amps in orig.items() if abs(amp_ratio[min(amps)] - 1.0) > 1e-10}
    amp_free = list(amp_chain)
    nxt = 0
    ac_free: Dict[int, List[int]] = {}
    for ck in sorted(orig):
        if ck in folded:
            continue
        amps = sorted(orig[ck])
        ac = amp_chain[amps[0]]
        for a in amps:
            amp_free[a] = nxt
            ac_free.setdefault(ac, []).append(nxt)
            nxt += 1
    for ck in sorted(orig):
        if ck not in folded:
            continue
        amps = sorted(orig[ck])
        fi = ac_free.get(amp_chain[amps[0]], [])
        for pos, a in enumerate(amps):
            if pos < len(fi):
                amp_free[a] = fi[pos]
    return amp_free, amp_ratio, nxt


def _apply_var_equal(cfg: ConfigParser, blocks, theta) -> None:
    """``Constraints.var_equal`` groups of parameter names share one slot: the first name
    of a group that has a slot owns it, the other members' slots are removed and their
    uses point at the owner (src/AmpGen.cu:2243-2303)."""
    if not cfg.var_equal or not theta:
        return
    name_to_slot = {t.name: j for j, t in enumerate(theta)}
    remap: Dict[int, int] = {}
    for group in cfg.var_equal:
        if len(group) < 2:
            continue
        owner = next((name_to_slot[n] for n in group if n in name_to_slot), None)
        if owner is None or owner in remap:
            continue
        for n in group:
            j = name_to_slot.get(n)
            if j is None or j == owner or j in remap:
                continue
            remap[j] = owner
    if not remap:
        return
    keep = [j for j in range(len(theta)) if j not in remap]
    new_index = {j: i for i, j in enumerate(keep)}
    for b in blocks:
        for inst in b.res:
            inst.slots = [(-1 if s < 0 else new_index[remap.get(s, s)]) for s in inst.slots]
    theta[:] = [theta[j] for j in keep]


def _open_table(cfg: ConfigParser, name: str, path: str):
    """A lineshape table file: as given (relative to the working directory, as the
    reference opens it), else relative to the config file's directory."""
    for cand in (path, os.path.join(cfg.base_dir, path)):
        if os.path.isfile(cand):
            return open(cand)
    raise ConfigError("%s: cannot open table file %s" % (name, path))


def _table(cfg: ConfigParser, rconf: ResonanceConfig, model: str) -> List[float]:
    """Shape table of a Hist or Interp lineshape (src/Resonance.cu:76-188).

    Hist: ``file`` (whitespace-separated values), ``bins`` (their number, >= 2) and
    ``range`` [lo, hi] -> [lo, hi, n, values...].  Interp: ``file`` with rows x Re Im
    -> [1, N, x_0, dx, re..., im...] when consecutive x differ by dx within 1e-10,
    else [-2, N, x..., re..., im...]; the interpolation is linear (the reference's
    configuration reader passes no method)."""
    opt = rconf.options
    if model == "Hist":
        if not opt.get("file") or not opt.get("bins") or not opt.get("range"):
            raise ConfigError("Hist %s requires file, bins and range" % rconf.name)
        try:
            n = int(opt["bins"])
        except ValueError:
            raise ConfigError("Hist %s: bins must be an integer" % rconf.name)
        if n < 2:
            raise ConfigError("Hist %s needs bins >= 2" % rconf.name)
        rng = []
        for tok in opt["range"].split(","):
            try:
                rng.append(float(tok))
            except ValueError:
                pass
        if len(rng) != 2 or not rng[1] > rng[0]:
            raise ConfigError("Hist %s needs range lo, hi with hi > lo" % rconf.name)
        with _open_table(cfg, rconf.name, opt["file"]) as f:
            vals = []
            for tok in f.read().split():
                try:
                    vals.append(float(tok))
                except ValueError:
                    break
        if len(vals) != n:
            raise ConfigError("Hist %s: file %s has %d values, expected %d" % (rconf.name, opt["file"], len(vals), n))
        return [rng[0], rng[1], float(n)] + vals
    if not opt.get("file"):
        raise ConfigError("Interp %s requires a file" % rconf.name)
    with _open_table(cfg, rconf.name, opt["file"]) as f:
        toks = f.read().split()
    x, re_, im_ = [], [], []
    for k in range(0, len(toks) - 2, 3):
        try:
            a, b, c = float(toks[k]), float(toks[k + 1]), float(toks[k + 2])
        except ValueError:
            break
        x.append(a)
        re_.append(b)
        im_.append(c)
    if len(x) < 2:
        raise ConfigError("Interp %s: file %s needs at least two rows x Re Im" % (rconf.name, opt["file"]))
    dx = x[1] - x[0]
    uniform = all(abs(x[i] - x[i - 1] - dx) <= 1e-10 for i in range(1, len(x)))
    if uniform:
        if dx == 0.0:
            raise ConfigError("Interp %s: repeated x values" % rconf.name)
        return [1.0, float(len(x)), x[0], dx] + re_ + im_
    return [-2.0, float(len(x))] + x + re_ + im_


def _chain_identical_groups(cfg: ConfigParser, chain) -> List[Tuple[List[str], bool]]:
    """Identical groups with at least two members among the chain's particles, members
    in Particles order, groups in the order of their names (src/Analysis.cu:4553-4580)."""
    in_chain = set()
    for st in chain.decay_steps:
        in_chain.add(st.mother)
        in_chain.update(st.daughters)
    members: Dict[str, List[str]] = {}
    boson: Dict[str, bool] = {}
    for p in cfg.particles:
        if p.identical_group and p.name in in_chain:
            members.setdefault(p.identical_group, []).append(p.name)
            boson[p.identical_group] = not p.is_fermion()
    return [(members[g], boson[g]) for g in sorted(members) if len(members[g]) >= 2]


def _identical_permutations(steps, root: str, groups, final_order: List[str]):
    """The permuted topologies of one chain (src/AmpGen.cu:194-403).

    For every element g != 1 of a group's permutation group (all of S_n, stabiliser
    swaps of sibling members included), the subtree of member g[i] takes the place of
    member i's, path by path.  The sign is the parity of g for fermions, +1 for bosons.
    Returns [(read, sign)] with ``read[f]`` the final-state index whose momentum the
    permuted topology reads where the chain names final particle f.

    Several groups combine the way the reference combines them: each element is a
    complete particle -> slot map and the earlier groups' map is laid over the later
    group's (src/AmpGen.cu:366-385), so only the first group's permutations survive,
    each repeated once per non-identity element of every later group, with the
    product of the signs."""
    maps: List[Dict[str, str]] = []
    signs: List[float] = []
    names = set()
    for st in steps:
        names.add(st.mother)
        names.update(st.daughters)
    for members, is_boson in groups:
        subs = [m for m in members if m != root and m in names]
        if len(subs) < 2:
            continue
        paths: List[Dict[Tuple[int, ...], str]] = []
        name_path: List[Dict[str, Tuple[int, ...]]] = []
        for m in subs:
            pth = {(): m}
            npth = {m: ()}
            changed = True
            while changed:
                changed = False
                for st in steps:
                    if st.mother not in npth:
                        continue
                    for side, d in ((1, st.daughters[0]), (2, st.daughters[1])):
                        q = npth[st.mother] + (side,)
                        if q not in pth:
                            pth[q] = d
                            npth[d] = q
                            changed = True
            paths.append(pth)
            name_path.append(npth)
        if any(set(pth) != set(paths[0]) for pth in paths):
            warnings.warn("identical group members have different decay structures (%s); "
                          "no symmetrisation" % ", ".join(subs))
            break
        if any(i != j and subs[j] in name_path[i] for i in range(len(subs)) for j in range(len(subs))):
            warnings.warn("identical group members nested inside each other (%s); "
                          "no symmetrisation" % ", ".join(subs))
            break
        n = len(subs)
        gmaps: List[Dict[str, str]] = []
        gsigns: List[float] = []
        for g in itertools.permutations(range(n)):
            if list(g) == list(range(n)):
                continue
            inv = sum(1 for i in range(n) for j in range(i + 1, n) if g[i] > g[j])
            sg = 1.0 if (is_boson or inv % 2 == 0) else -1.0
            # placement: the momentum of particle Y goes to the slot of particle X
            m = {nm: nm for nm in names}
            for i in range(n):
                for q, nm_i in paths[i].items():
                    m[paths[g[i]][q]] = nm_i
            gmaps.append(m)
            gsigns.append(sg)
        if not maps:
            maps, signs = gmaps, gsigns
        else:
            merged, msigns = [], []
            for a in range(len(maps)):
                for b in range(len(gmaps)):
                    mm = dict(gmaps[b])
                    mm.update(maps[a])
                    merged.append(mm)
                    msigns.append(signs[a] * gsigns[b])
            maps, signs = merged, msigns
    out = []
    for m, sg in zip(maps, signs):
        read = list(range(len(final_order)))
        for y, x in m.items():
            if x in final_order and y in final_order:
                read[final_order.index(x)] = final_order.index(y)
        out.append((read, sg))
    return out


def _custom_expr(text: str, names: List[str], L_values: List[int], nd) -> List[str]:
    """Parse a Custom expression at node ``nd`` whose waves take the given L values.  A
    daughter's mass is its configured mass when positive, 0 for a final-state particle
    configured massless (a photon), as in the built-in models' q0."""
    def dmass(m, ref):
        return m if (m > 0 or (m == 0 and ref >= 0)) else None
    ctx = customexpr.Context(names, L_values, nd.bf_d, dmass(nd.masses[1], nd.d1_ref),
                             dmass(nd.masses[2], nd.d2_ref))
    return customexpr.serialize(customexpr.parse(text, ctx))


def _symmetrise(topologies: List[Topology], blocks: List[Block], perms: Dict[int, list],
                amp_names: List[str], amp_free: List[int], amp_ratio: List[float]) -> int:
    """Append one permuted copy of every topology with identical particles and of each of
    its blocks: the copy reads the permuted momenta, shares the block's resonance
    parameters, and its amplitudes take the original's coupling times the permutation's
    sign, so that sum_a v_a A_a carries A(x) + sum_sigma sgn(sigma) A(x_sigma)
    (src/AmpGen.cu:1440-1460).  Returns the new amplitude count."""
    n_amp = len(amp_names)
    n_orig = len(blocks)
    for b in range(n_orig):
        if blocks[b].wave < 0:
            blocks[b].wave = b
    copies: Dict[Tuple[int, int], int] = {}
    for b in range(n_orig):
        blk = blocks[b]
        for k, (read, sg) in enumerate(perms.get(blk.topo, []), start=1):
            key = (blk.topo, k)
            if key not in copies:
                t = topologies[blk.topo]
                nodes = []
                for nd in t.nodes:
                    c = copy.copy(nd)
                    c.d1_ref = read[nd.d1_ref] if nd.d1_ref >= 0 else nd.d1_ref
                    c.d2_ref = read[nd.d2_ref] if nd.d2_ref >= 0 else nd.d2_ref
                    nodes.append(c)
                copies[key] = len(topologies)
                topologies.append(Topology(t.chain, list(t.comb), nodes, list(t.sl_combos)))
            blocks.append(Block(copies[key], n_amp, blk.nsl, copy.deepcopy(blk.res), blk.chain_key,
                                wave=blk.wave, sigma=k, c0=blk.c0, src=b))
            for c in range(blk.nsl):
                a = blk.site + c
                amp_names.append("%s#perm%d" % (amp_names[a], k))
                amp_free.append(amp_free[a])
                amp_ratio.append(amp_ratio[a] * sg)
            n_amp += blk.nsl
    return n_amp


def build_plan(cfg: ConfigParser, fit_mode: int = 1) -> Plan:
    if fit_mode != 1:
        raise ConfigError("only fit mode 1 (vspace) is built; chain x step mode is not")
    particles = cfg.particles
    pmap = {p.name: p for p in particles}
    final_order = list(cfg.data_order)
    if not final_order:
        raise ConfigError("Data.order is required")
    for nm in final_order:
        if nm not in pmap:
            raise ConfigError("Data.order particle '%s' is not in Particles" % nm)
    npol, polmap = _polarization(particles)

    topologies: List[Topology] = []
    blocks: List[Block] = []
    theta: List[ThetaSlot] = []
    owners: Dict[str, Tuple[int, int]] = {}
    amp_names: List[str] = []
    chain_keys: List[str] = []
    amp_chain_key: List[str] = []
    gls = 0
    initial: Optional[str] = None
    perms: Dict[int, list] = {}          # topology -> identical-particle permutations

    for chain in cfg.decay_chains:
        # intermediate_resonance_map (src/Info.cu:60-120)
        ir_map: Dict[Tuple[str, int, int], List[ResonanceConfig]] = {}
        inter_particles: List[List[Tuple[str, int, int]]] = []
        for rc in chain.resonance_chains:
            jps = []
            for sc in rc.spin_chains:
                if len(sc.spin_parity) < 2:
                    raise ConfigError("intermediate %s needs [J, P]" % rc.intermediate)
                key = (rc.intermediate, sc.spin_parity[0], sc.spin_parity[1])
                lst = []
                for rname in sc.resonances:
                    r = cfg.resonances.get(rname)
                    if r is not None and (r.J < 0 or r.J == sc.spin_parity[0]):
                        lst.append(r)
                ir_map[key] = lst
                jps.append(key)
            inter_particles.append(jps)
        combs = [list(c) for c in itertools.product(*inter_particles)] if inter_particles else [[]]

        for comb in combs:
            jp = {name: (s, p) for (name, s, p) in comb}
            nodes: List[Node] = []
            for st in chain.decay_steps:
                if len(st.daughters) != 2:
                    raise ConfigError("decay steps must be two-body")

                def qn(nm):
                    if nm in jp:
                        return jp[nm]
                    if nm in pmap:
                        return pmap[nm].spin, pmap[nm].parity
                    return 0, 0
                (s0, p0), (s1, p1), (s2, p2) = qn(st.mother), qn(st.daughters[0]), qn(st.daughters[1])
                if 0 in (s0, s1, s2):
                    raise ConfigError("unresolved quantum numbers in chain %s" % chain.name)
                # identical daughters: (-1)^(L+S) = +1 (bosons) or -1 (fermions)
                # (src/Info.cu:150-162)
                g1 = pmap[st.daughters[0]].identical_group if st.daughters[0] in pmap else ""
                g2 = pmap[st.daughters[1]].identical_group if st.daughters[1] in pmap else ""
                ident = bool(g1) and g1 == g2
                boson = not pmap[st.daughters[0]].is_fermion() if ident else True
                sl = com_sl((s0, s1, s2), (p0, p1, p2), st.p_break, ident, boson,
                            cfg.global_maxL, st.sl_filter)
                masses = tuple(pmap[nm].mass if nm in pmap else -1.0
                               for nm in (st.mother, st.daughters[0], st.daughters[1]))
                nodes.append(Node(st.mother, st.daughters[0], st.daughters[1], (s0, s1, s2),
                                  (p0, p1, p2), sl, st.has_bf, st.bf_d, masses))
            # momentum references
            root = nodes[0].mother
            if initial is None:
                initial = root
            elif initial != root:
                raise ConfigError("all chains must start from the same initial particle")
            mothers = {nd.mother: k for k, nd in enumerate(nodes)}
            for nd in nodes:
                refs = []
                for d in (nd.d1, nd.d2):
                    if d in final_order:
                        refs.append(final_order.index(d))
                    elif d in mothers and mothers[d] > 0:
                        refs.append(-(mothers[d] + 1))
                    else:
                        raise ConfigError("particle '%s' is neither final nor decaying" % d)
                nd.d1_ref, nd.d2_ref = refs
                nd.lmin = min((L for _, L in nd.sl_list), default=0)
            sl_combos = [tuple(c) for c in itertools.product(*[nd.sl_list for nd in nodes])]
            topo_idx = len(topologies)
            topologies.append(Topology(chain.name, comb, nodes, sl_combos))
            groups = _chain_identical_groups(cfg, chain)
            if groups:
                perms[topo_idx] = _identical_permutations(chain.decay_steps, nodes[0].mother,
                                                          groups, final_order)

            # resonance combinations (last intermediate fastest)
            res_lists = [ir_map[key] for key in comb]
            for rcombo in (itertools.product(*res_lists) if res_lists else [()]):
                insts: List[ResInst] = []
                per_wave: List[tuple] = []    # Custom resonances whose expression takes each wave's L
                free_lists: List[List[int]] = []
                range_lists: List[List[List[float]]] = []
                for key, rconf in zip(comb, rcombo):
                    model = canonical_model(rconf.type)
                    if model == "Custom":
                        vals = [float(x) for x in rconf.parameters]
                        pn = rconf.options.get("params")
                        names = pn.split(",") if pn else ["p%d" % i for i in range(len(vals))]
                        if len(names) != len(vals):
                            raise ConfigError("Custom %s: %d paramNames for %d paramValues"
                                              % (rconf.name, len(names), len(vals)))
                    else:
                        vals, names = ordered_params(model, rconf.parameters)
                    # Constraints.fix_var overrides values (src/Analysis.cu:4632-4650)
                    for pi, pn in enumerate(names):
                        nm = rconf.name + "_" + pn
                        if nm in cfg.fix_var and nm not in cfg.free_var:
                            vals[pi] = cfg.fix_var[nm]
                    node_k = mothers.get(key[0], -1)
                    if node_k <= 0:
                        raise ConfigError("intermediate %s does not decay in chain %s"
                                          % (key[0], chain.name))
                    if model == "Flatte" and len(rconf.channels) < len(vals) - 1:
                        raise ConfigError("Flatte %s needs one channel per coupling" % rconf.name)
                    chans = [(float(c[0]), float(c[1])) for c in rconf.channels]
                    inst = ResInst(rconf.name, key[0], model, vals, names, chans, node_k, [-1] * len(vals))
                    if model in ("Hist", "Interp"):
                        inst.table = _table(cfg, rconf, model)
                    if model == "Custom":
                        text = rconf.options.get("expression", rconf.options.get("expr"))
                        if not text:
                            raise ConfigError("Custom %s has no expression" % rconf.name)
                        try:
                            inst.expr = _custom_expr(text, names, [c[node_k][1] for c in sl_combos],
                                                     nodes[node_k])
                        except PerWaveL:
                            per_wave.append((len(insts), text, names, node_k))
                    insts.append(inst)
                    if rconf.free:
                        if len(rconf.free) == 1 and rconf.free[0] == -1:
                            fidx = list(range(len(vals)))
                        else:
                            fidx = [i for i in rconf.free if 0 <= i < len(vals)]
                        fidx = [i for i in fidx
                                if not (rconf.name + "_" + names[i] in cfg.fix_var
                                        and rconf.name + "_" + names[i] not in cfg.free_var)]
                        free_lists.append(fidx)
                    else:
                        free_lists.append([])
                    range_lists.append(rconf.free_range)

                # theta slot registration (src/AmpGen.cu:2165-2242)
                for inst, fidx, rng in zip(insts, free_lists, range_lists):
                    own = owners.get(inst.name)
                    if own is not None and fidx:
                        ob, oi = own
                        src = blocks[ob].res[oi] if ob < len(blocks) else insts[oi]
                        inst.slots = list(src.slots)
                        continue
                    owners[inst.name] = (len(blocks), insts.index(inst))
                    for si, p in enumerate(fidx):
                        val = inst.params[p]
                        if si < len(rng) and len(rng[si]) >= 2:
                            lo, hi = rng[si][0], rng[si][1]
                        else:
                            half = abs(val) * 0.5
                            lo, hi = val - half, val + half
                        nm = inst.name + "_" + inst.param_names[p]
                        if nm in cfg.var_range:
                            lo, hi = cfg.var_range[nm]
                        inst.slots[p] = len(theta)
                        theta.append(ThetaSlot(nm, inst.name, p, val, lo, hi))

                tagmap = {inst.tag: inst.name for inst in insts}
                chain_key = chain.name + "".join("_%s[%s]" % (i.tag, i.name) for i in insts)
                if not per_wave:
                    blocks.append(Block(topo_idx, gls, len(sl_combos), insts, chain_key))
                else:
                    # one block per run of SL combinations with one L at every such node: each
                    # run's expressions take that L; the runs form one partial wave
                    lkey = [tuple(c[k][1] for (_, _, _, k) in per_wave) for c in sl_combos]
                    first, c0 = len(blocks), 0
                    while c0 < len(sl_combos):
                        c1 = c0 + 1
                        while c1 < len(sl_combos) and lkey[c1] == lkey[c0]:
                            c1 += 1
                        sub = [copy.copy(x) for x in insts]
                        for (ii, text, nms, k) in per_wave:
                            sub[ii].expr = _custom_expr(text, nms, [sl_combos[c0][k][1]], nodes[k])
                        blocks.append(Block(topo_idx, gls + c0, c1 - c0, sub, chain_key, wave=first, c0=c0))
                        c0 = c1
                for c in sl_combos:
                    parts = []
                    for nd, (S, L) in zip(nodes, c):
                        parts.append("%s→%s+%s_SL(%d,%d)" % (
                            tagmap.get(nd.mother, nd.mother), tagmap.get(nd.d1, nd.d1),
                            tagmap.get(nd.d2, nd.d2), S, L))
                    amp_names.append("_".join(parts))
                    amp_chain_key.append(chain_key)
                    if chain_key not in chain_keys:
                        chain_keys.append(chain_key)
                gls += len(sl_combos)

    n_amp = gls
    _apply_var_equal(cfg, blocks, theta)
    amp_free, amp_ratio, n_free = _vspace_map(cfg, amp_chain_key, chain_keys)
    free_names: Dict[int, str] = {}
    for a in range(n_amp):
        if abs(amp_ratio[a] - 1.0) < 1e-10:
            free_names[amp_free[a]] = amp_names[a]
    param_names = [free_names.get(i, "") for i in range(n_free)] + [t.name for t in theta]
    n_amp = _symmetrise(topologies, blocks, perms, amp_names, amp_free, amp_ratio)

    gauss = []
    for nm, sig in cfg.gauss_constr.items():
        for j, t in enumerate(theta):
            if t.name == nm:
                gauss.append((j, t.init, float(sig)))

    return Plan(particles=particles, final_order=final_order, initial=initial or "",
                npol=npol, polmap=polmap, topologies=topologies, blocks=blocks,
                n_amp=n_amp, amp_names=amp_names, theta=theta, amp_free=amp_free,
                amp_ratio=amp_ratio, n_free=n_free, param_names=param_names, gauss=gauss)
