"""config.yml schema reader for the likelihood path.

Mirrors the semantics of the reference's ``ConfigParser`` (src/Config.cu:138-1310,
include/Config.cuh:10-190) for the sections the likelihood needs: Particles, Data,
DecayChains (compact and expanded forms), Resonances (inline or external file),
Constraints and ``precision``.  The reference parses with yaml-cpp; here PyYAML is
used with a loader that accepts the schema's sequence-valued mapping keys
(``- [J: 1, P: -1]: [rho770]``), which PyYAML's stock loader rejects.

Plot/legend sections are accepted and ignored: they belong to post-fit output.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field
from typing import Any, Dict, List, Optional, Tuple

import yaml


class ConfigError(ValueError):
    """Raised for a config the likelihood path cannot represent."""


# --------------------------------------------------------------------------
# YAML loading
# --------------------------------------------------------------------------
class _Key(tuple):
    """Hashable stand-in for a sequence used as a mapping key."""


def _freeze(obj):
    if isinstance(obj, list):
        return _Key(_freeze(x) for x in obj)
    if isinstance(obj, dict):
        return _Key(tuple((k, _freeze(v)) for k, v in obj.items()))
    return obj


class _Loader(yaml.SafeLoader):
    pass


def _construct_mapping(loader, node, deep=False):
    # Keep insertion order (yaml-cpp iterates a map in document order).
    loader.flatten_mapping(node)
    out = {}
    for key_node, value_node in node.value:
        key = loader.construct_object(key_node, deep=True)
        if isinstance(key, (list, dict)):
            key = _freeze(key)
        out[key] = loader.construct_object(value_node, deep=deep)
    return out


_Loader.add_constructor(yaml.resolver.BaseResolver.DEFAULT_MAPPING_TAG, _construct_mapping)


# top-level sections the likelihood does not read; yaml-cpp accepts some forms in them
# that strict YAML rejects (e.g. ``bins: 100; range: [1.0, 2.6]`` in Plot)
_UNUSED_SECTIONS = ("Plot",)


def _drop_sections(text: str, names) -> str:
    out, skip = [], False
    for line in text.splitlines():
        top = line and not line[0].isspace() and not line.startswith("#")
        if top:
            key = line.split(":", 1)[0].strip()
            skip = key in names
        if not skip:
            out.append(line)
    return "\n".join(out) + "\n"


def parse_yaml_text(text: str):
    """YAML of a config; on a syntax error the sections the likelihood does not use are
    dropped and the text parsed again."""
    try:
        return yaml.load(text, Loader=_Loader)
    except yaml.YAMLError:
        return yaml.load(_drop_sections(text, _UNUSED_SECTIONS), Loader=_Loader)


def load_yaml(path: str):
    with open(path, "r", encoding="utf-8") as f:
        return parse_yaml_text(f.read())


def trans_j_value(s) -> float:
    """Spin string -> J ("1/2" -> 0.5, "1" -> 1.0).  src/Config.cu transJValue."""
    s = str(s).strip()
    if "/" in s:
        a, b = s.split("/")
        return float(a) / float(b)
    return float(s)


def _jp_from_key(key) -> List[int]:
    """``[J: 1, P: -1]`` key -> [2J+1, P]  (src/Config.cu:700-715)."""
    sp: List[int] = []
    items = key if isinstance(key, (tuple, list)) else [key]
    for it in items:
        if isinstance(it, tuple):
            d = dict(it)
        elif isinstance(it, dict):
            d = it
        else:
            continue
        if "J" in d:
            sp.append(int(round(2 * trans_j_value(d["J"]) + 1)))
        if "P" in d:
            sp.append(int(d["P"]))
    return sp


# --------------------------------------------------------------------------
# Schema records (include/Config.cuh:10-75)
# --------------------------------------------------------------------------
@dataclass
class Particle:
    name: str
    spin: int            # 2J+1
    parity: int
    mass: float
    identical_group: str = ""
    polarization_2m: List[int] = field(default_factory=list)

    def is_fermion(self) -> bool:
        return (self.spin - 1) % 2 != 0

    def is_polarized(self) -> bool:
        return bool(self.polarization_2m)


@dataclass
class ResonanceConfig:
    name: str
    J: int               # 2J+1, -1 when not given
    P: int
    type: str
    parameters: List[float]
    channels: List[List[float]] = field(default_factory=list)
    free: List[int] = field(default_factory=list)
    free_range: List[List[float]] = field(default_factory=list)
    options: Dict[str, str] = field(default_factory=dict)
    has_bf: bool = True
    has_bf_explicit: bool = False
    bf_d: float = math.nan


@dataclass
class SpinChain:
    spin_parity: List[int]
    resonances: List[str]


@dataclass
class ResonanceChain:
    intermediate: str
    spin_chains: List[SpinChain]


@dataclass
class DecayStep:
    mother: str
    daughters: List[str]
    has_bf: bool = True
    has_bf_explicit: bool = False
    bf_d: float = math.nan
    p_break: bool = False
    sl_filter: List[Tuple[int, int]] = field(default_factory=list)


@dataclass
class DecayChain:
    name: str
    decay_steps: List[DecayStep] = field(default_factory=list)
    resonance_chains: List[ResonanceChain] = field(default_factory=list)


@dataclass
class Constraint:
    type: str
    names: List[str]
    values: List[complex]


def _sl_rows(raw) -> List[Tuple[int, int]]:
    """``sl`` whitelist: flat [S, L] or nested; S physical -> 2S+1 (src/Config.cu:633-646)."""
    if not raw:
        return []
    rows = raw if isinstance(raw[0], (list, tuple)) else [raw]
    out = []
    for r in rows:
        if len(r) >= 2:
            out.append((int(round(2.0 * float(r[0]) + 1.0)), int(r[1])))
    return out


class ConfigParser:
    """Reads a reference-schema config.yml.  Attribute names follow the reference getters."""

    def __init__(self, config_file: str = "config.yml", text: Optional[str] = None):
        self.config_file = config_file
        self.base_dir = os.path.dirname(os.path.abspath(config_file)) if config_file else "."
        self.particles: List[Particle] = []
        self.decay_chains: List[DecayChain] = []
        self.resonances: Dict[str, ResonanceConfig] = {}
        self.data_files: Dict[str, List[str]] = {}
        self.data_order: List[str] = []
        self.constraints: List[Constraint] = []
        self.global_maxL = -1
        self.global_bf_d = 3.0
        self.global_has_bf = True
        self.precision = "auto"
        self.fix_var: Dict[str, float] = {}
        self.free_var: set = set()
        self.var_range: Dict[str, Tuple[float, float]] = {}
        self.var_equal: List[List[str]] = []
        self.gauss_constr: Dict[str, float] = {}
        self._chain_filter: List[str] = []

        if text is not None:
            cfg = parse_yaml_text(text)
        else:
            cfg = load_yaml(config_file)
        if not isinstance(cfg, dict):
            raise ConfigError("config root must be a mapping")
        self._parse(cfg)

    # src/Config.cu:138-248
    def _parse(self, cfg):
        if "Particles" in cfg:
            self._parse_particles(cfg["Particles"])
        if "Data" in cfg:
            self._parse_data(cfg["Data"])
        cons = cfg.get("Constraints") or {}
        if isinstance(cons, dict) and cons.get("identical"):
            self._parse_identical(cons["identical"])
        if isinstance(cons, dict) and "chains" in cons:
            ch = cons["chains"]
            if isinstance(ch, str):
                path = os.path.join(self.base_dir, ch)
                if os.path.exists(path):
                    with open(path) as f:
                        for line in f:
                            pat = line.strip()
                            if pat and not pat.startswith("#"):
                                self._chain_filter.append(pat)
            elif isinstance(ch, list):
                self._chain_filter = [str(x) for x in ch]
        if "Resonances" in cfg:
            res = cfg["Resonances"]
            if isinstance(res, str):
                node = load_yaml(os.path.join(self.base_dir, res))
                if isinstance(node, dict) and "Resonances" in node:
                    node = node["Resonances"]
                self._parse_resonances(node or {})
            else:
                self._parse_resonances(res or {})
        if "DecayChains" in cfg:
            self._parse_decay_chains(cfg["DecayChains"])
        if isinstance(cons, dict):
            self._parse_constraints(cons)
        self._resolve_step_bf()
        if "precision" in cfg:
            self.precision = str(cfg["precision"])
        if self._chain_filter and self.decay_chains:
            self.decay_chains = [dc for dc in self.decay_chains
                                 if any(p in dc.name for p in self._chain_filter)]

    def is_valid(self) -> bool:
        return bool(self.particles)

    # src/Config.cu:477-521
    def _parse_particles(self, node):
        for name, props in node.items():
            spin = int(round(2 * trans_j_value(props["J"]) + 1))
            p = Particle(name=str(name), spin=spin, parity=int(props["P"]),
                         mass=float(props["mass"]))
            pol = props.get("polarization")
            if isinstance(pol, list):
                p.polarization_2m = [int(round(2 * trans_j_value(m))) for m in pol]
            self.particles.append(p)

    # src/Config.cu:523-552
    def _parse_data(self, node):
        if "order" in node:
            self.data_order = [str(x) for x in node["order"]]
        for key in ("data", "phsp", "phsp_truth", "bkg", "bkg_weights",
                    "data_weights", "phsp_weights"):
            if key in node:
                v = node[key]
                self.data_files[key] = [str(x) for x in (v if isinstance(v, list) else [v])]

    def _parse_identical(self, groups):
        for gi, group in enumerate(groups, start=1):
            gname = "identical%d" % gi
            for p in self.particles:
                if p.name in group:
                    p.identical_group = gname

    # src/Config.cu:1041-1138
    def _parse_resonances(self, node):
        for name, props in node.items():
            J = int(round(2 * trans_j_value(props["J"]) + 1)) if "J" in props else -1
            P = int(props["P"]) if "P" in props else 0
            params = props.get("paramValues", props.get("parameters", [])) or []
            r = ResonanceConfig(name=str(name), J=J, P=P, type=str(props["model"]),
                                parameters=[float(x) for x in params])
            if "channels" in props:
                r.channels = [[float(a) for a in ch] for ch in props["channels"]]
            if "free" in props:
                r.free = [int(x) for x in props["free"]]
            if "free_range" in props:
                r.free_range = [[float(a) for a in rg] for rg in props["free_range"]]
            for k in ("file", "bins", "extrapolate", "expr", "expression"):
                if k in props:
                    r.options[k] = str(props[k])
            if "range" in props:
                # the reference joins the values with std::to_string ("%f")
                r.options["range"] = ",".join("%f" % float(x) for x in props["range"])
            pn = props.get("paramNames", props.get("params"))
            if pn:
                r.options["params"] = ",".join(str(x) for x in pn)
            if "has_bf" in props:
                r.has_bf = bool(props["has_bf"])
                r.has_bf_explicit = True
            if "bf_d" in props:
                r.bf_d = float(props["bf_d"])
            self.resonances[r.name] = r

    # src/Config.cu:554-1039
    def _parse_decay_chains(self, node):
        particle_names = {p.name for p in self.particles}
        for chain_name, data in node.items():
            chain_name = str(chain_name)
            mother = None
            if "decay" not in data:
                for key, val in data.items():
                    if key in ("intermediates", "legend", "legends", "symmetrize"):
                        continue
                    if key in particle_names and isinstance(val, list) and val \
                            and isinstance(val[0], list) and len(val[0]) >= 2:
                        mother = key
                        break
            if mother is not None:
                self._expand_compact(chain_name, mother, data)
            else:
                self._parse_expanded(chain_name, data)

    def _parse_int_mode(self, node) -> dict:
        mode = dict(d1=str(node[0]), d2=str(node[1]), has_bf=True, has_bf_explicit=False,
                    bf_d=math.nan, p_break=False, sl=[])
        if len(node) >= 3 and isinstance(node[2], dict):
            o = node[2]
            if "has_bf" in o:
                mode["has_bf"] = bool(o["has_bf"])
                mode["has_bf_explicit"] = True
            if "bf_d" in o:
                mode["bf_d"] = float(o["bf_d"])
            if "p_break" in o:
                mode["p_break"] = bool(o["p_break"])
            if "sl" in o:
                mode["sl"] = _sl_rows(o["sl"])
        return mode

    def _res_chain_from_spec(self, int_name, spec) -> ResonanceChain:
        rc = ResonanceChain(intermediate=int_name, spin_chains=[])
        for spin_node in spec:
            if isinstance(spin_node, dict):
                for key, names in spin_node.items():
                    rc.spin_chains.append(SpinChain(_jp_from_key(key), [str(n) for n in names]))
            elif isinstance(spin_node, list):
                # auto-detect: group by (J, P) of the Resonances section, ordered by (J, P)
                groups: Dict[Tuple[int, int], List[str]] = {}
                for rname in spin_node:
                    r = self.resonances.get(str(rname))
                    if r is None:
                        continue
                    if r.J < 0:
                        raise ConfigError("resonance '%s' has no J/P; use the explicit "
                                          "[J: .., P: ..] form" % rname)
                    groups.setdefault((r.J, r.P), []).append(str(rname))
                for jp in sorted(groups):
                    rc.spin_chains.append(SpinChain([jp[0], jp[1]], groups[jp]))
            elif isinstance(spin_node, str):
                # a bare name in a flat list: treated as its own auto-detect group
                r = self.resonances.get(spin_node)
                if r is not None and r.J >= 0:
                    rc.spin_chains.append(SpinChain([r.J, r.P], [spin_node]))
        return rc

    def _expand_compact(self, chain_name, mother, data):
        channels = data[mother]
        int_modes: Dict[str, List[dict]] = {}
        for key, val in data.items():
            if key in (mother, "intermediates", "legend", "legends", "symmetrize"):
                continue
            if not isinstance(val, list) or not val:
                continue
            if isinstance(val[0], list):
                int_modes[key] = [self._parse_int_mode(m) for m in val]
            elif len(val) >= 2:
                int_modes[key] = [self._parse_int_mode(val)]

        def get_mode(name, idx=0):
            m = int_modes.get(name)
            if m is None or idx >= len(m):
                return None
            return m[idx]

        res_chain_map: Dict[str, ResonanceChain] = {}
        inter = data.get("intermediates") or {}
        for int_name, spec in inter.items():
            res_chain_map[str(int_name)] = self._res_chain_from_spec(str(int_name), spec)

        for ch in channels:
            bachelor, intermediate = str(ch[0]), str(ch[1])
            has_bf1 = has_bf2 = True
            exp1 = exp2 = False
            bfd1 = bfd2 = math.nan
            pb1 = pb2 = False
            has_ch_pb = False
            sl1: List[Tuple[int, int]] = []
            if len(ch) >= 3 and isinstance(ch[2], dict):
                o = ch[2]
                if "has_bf" in o:
                    v = o["has_bf"]
                    if isinstance(v, list):
                        has_bf1, has_bf2 = bool(v[0]), bool(v[1])
                    else:
                        has_bf1 = has_bf2 = bool(v)
                    exp1 = exp2 = True
                if "bf_d" in o:
                    v = o["bf_d"]
                    if isinstance(v, list):
                        bfd1, bfd2 = float(v[0]), float(v[1])
                    else:
                        bfd1 = bfd2 = float(v)
                if "p_break" in o:
                    has_ch_pb = True
                    v = o["p_break"]
                    if isinstance(v, list):
                        pb1, pb2 = bool(v[0]), bool(v[1])
                    else:
                        pb1 = pb2 = bool(v)
                if "sl" in o:
                    sl1 = _sl_rows(o["sl"])

            n_modes, bachelor_drives = 1, False
            if intermediate in int_modes:
                n_modes = len(int_modes[intermediate])
            elif bachelor in int_modes:
                n_modes = len(int_modes[bachelor])
                bachelor_drives = True

            for mi in range(n_modes):
                if n_modes > 1:
                    m = get_mode(bachelor if bachelor_drives else intermediate, mi)
                    other = intermediate if bachelor_drives else bachelor
                    if m and (m["d1"] == other or m["d2"] == other):
                        continue
                chain = DecayChain(name=chain_name + "_" + (bachelor if bachelor_drives else intermediate))
                if n_modes > 1:
                    chain.name += "_%d" % mi
                chain.decay_steps.append(DecayStep(mother, [bachelor, intermediate], has_bf1, exp1,
                                                   bfd1, pb1, list(sl1)))
                inter_decays = intermediate in int_modes
                bach_decays = bachelor in int_modes
                if not inter_decays and intermediate in res_chain_map:
                    chain.resonance_chains.append(res_chain_map[intermediate])
                if not bach_decays and bachelor in res_chain_map:
                    chain.resonance_chains.append(res_chain_map[bachelor])
                if not inter_decays and not bach_decays:
                    self.decay_chains.append(chain)
                    continue
                queue = []
                if inter_decays:
                    ifm = get_mode(intermediate, 0 if bachelor_drives else mi)
                    s2_bf = has_bf2 if exp2 else (ifm["has_bf"] if ifm and ifm["has_bf_explicit"] else True)
                    s2_exp = exp2 or bool(ifm and ifm["has_bf_explicit"])
                    s2_d = bfd2 if not math.isnan(bfd2) else (ifm["bf_d"] if ifm else math.nan)
                    s2_pb = pb2 if has_ch_pb else (ifm["p_break"] if ifm else False)
                    queue.append(dict(name=intermediate, first=not bachelor_drives, has_bf=s2_bf,
                                      exp=s2_exp, pb=s2_pb, bf_d=s2_d, sl=ifm["sl"] if ifm else []))
                if bach_decays:
                    bm = get_mode(bachelor, mi if bachelor_drives else 0)
                    queue.append(dict(name=bachelor, first=bachelor_drives,
                                      has_bf=bm["has_bf"] if bm and bm["has_bf_explicit"] else True,
                                      exp=bool(bm and bm["has_bf_explicit"]),
                                      pb=bm["p_break"] if bm else False,
                                      bf_d=bm["bf_d"] if bm else math.nan,
                                      sl=bm["sl"] if bm else []))
                while queue:
                    item = queue.pop(0)
                    mode = get_mode(item["name"], mi if item["first"] else 0)
                    if mode is None:
                        continue
                    chain.decay_steps.append(DecayStep(item["name"], [mode["d1"], mode["d2"]],
                                                       item["has_bf"], item["exp"], item["bf_d"],
                                                       item["pb"], list(item["sl"])))
                    if item["name"] in res_chain_map:
                        chain.resonance_chains.append(res_chain_map[item["name"]])
                    for d in (mode["d1"], mode["d2"]):
                        if d in int_modes:
                            dm = get_mode(d, 0)
                            queue.append(dict(name=d, first=False,
                                              has_bf=dm["has_bf"] if dm and dm["has_bf_explicit"] else True,
                                              exp=bool(dm and dm["has_bf_explicit"]),
                                              pb=dm["p_break"] if dm else False,
                                              bf_d=dm["bf_d"] if dm else math.nan,
                                              sl=dm["sl"] if dm else []))
                self.decay_chains.append(chain)

    def _parse_expanded(self, chain_name, data):
        chain = DecayChain(name=chain_name)
        for step_node in data.get("decay", []) or []:
            for m, d in step_node.items():
                st = DecayStep(mother=str(m), daughters=[])
                if isinstance(d, list):
                    st.daughters = [str(x) for x in d]
                elif isinstance(d, dict):
                    st.daughters = [str(x) for x in d["daughters"]]
                    if "has_bf" in d:
                        st.has_bf = bool(d["has_bf"])
                        st.has_bf_explicit = True
                    if "bf_d" in d:
                        st.bf_d = float(d["bf_d"])
                    if "p_break" in d:
                        st.p_break = bool(d["p_break"])
                    if "sl" in d:
                        raw = d["sl"]
                        rows = raw if isinstance(raw[0], list) else [raw]
                        st.sl_filter = [(int(r[0]), int(r[1])) for r in rows if len(r) >= 2]
                chain.decay_steps.append(st)
        for key, spec in data.items():
            if key in ("decay", "legend", "symmetrize"):
                continue
            rc = ResonanceChain(intermediate=str(key), spin_chains=[])
            for spin_node in spec or []:
                if isinstance(spin_node, dict):
                    for k2, names in spin_node.items():
                        rc.spin_chains.append(SpinChain(_jp_from_key(k2), [str(n) for n in names]))
            chain.resonance_chains.append(rc)
        self.decay_chains.append(chain)

    # src/Config.cu:1178-1310
    def _parse_constraints(self, node):
        if "maxL" in node:
            self.global_maxL = int(node["maxL"])
        if "bf_d" in node:
            self.global_bf_d = float(node["bf_d"])
        if "has_bf" in node:
            self.global_has_bf = bool(node["has_bf"])
        for k, v in (node.get("fix_var") or {}).items():
            self.fix_var[str(k)] = float(v)
        for k in node.get("free_var") or []:
            self.free_var.add(str(k))
        for k, v in (node.get("var_range") or {}).items():
            if isinstance(v, list) and len(v) >= 2:
                self.var_range[str(k)] = (float(v[0]), float(v[1]))
        for g in node.get("var_equal") or []:
            if len(g) >= 2:
                self.var_equal.append([str(x) for x in g])
        for k, v in (node.get("gauss_constr") or {}).items():
            self.gauss_constr[str(k)] = float(v)
        for item in node.get("trans") or []:
            for names, values in item.items():
                c = Constraint(type="trans", names=[str(x) for x in names], values=[])
                if isinstance(values, list):
                    if values and isinstance(values[0], list):
                        c.values = [complex(float(r[0]), float(r[1])) for r in values if len(r) >= 2]
                    else:
                        c.values = [complex(float(v), 0.0) for v in values]
                else:
                    c.values = [complex(float(values), 0.0)]
                self.constraints.append(c)

    # src/Config.cu:1140-1176
    def _resolve_step_bf(self):
        for chain in self.decay_chains:
            for step in chain.decay_steps:
                cfg = self.resonances.get(step.mother)
                if cfg is None:
                    for rc in chain.resonance_chains:
                        if rc.intermediate != step.mother or not rc.spin_chains:
                            continue
                        names = rc.spin_chains[0].resonances
                        if names and names[0] in self.resonances:
                            cfg = self.resonances[names[0]]
                        break
                if not step.has_bf_explicit:
                    step.has_bf = cfg.has_bf if (cfg is not None and cfg.has_bf_explicit) \
                        else self.global_has_bf
                if math.isnan(step.bf_d):
                    step.bf_d = cfg.bf_d if (cfg is not None and not math.isnan(cfg.bf_d)) \
                        else self.global_bf_d
