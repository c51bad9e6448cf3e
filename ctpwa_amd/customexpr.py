"""Custom lineshape expressions (``model: Custom``) -> prefix token list for the plan.

The expression language follows the reference's DSL (src/CustomExpr.cu:230-560):

* numbers, ``pi``, the imaginary unit ``1j`` / ``1i`` (and ``i``);
* variables ``m`` (invariant mass of the node), ``q`` (its breakup momentum), ``q0``
  (breakup at the first parameter, computeQ0AD), ``L`` (orbital angular momentum of the
  wave at the node), ``d`` (barrier radius), ``md1``/``md2`` (the q0 daughter masses);
* parameters by their ``paramNames`` (default ``p0, p1, ...``);
* ``+ - * /``, right-associative ``^`` (complex power exp(b log a)), unary minus;
* functions exp, log/ln, sin, cos, sqrt, csqrt (both principal complex square roots),
  abs, Re/real, Im/imag, conj;
* model functions BW(m0, w0), BWR(m0, g0[, L, d]) (L, d literals, defaults 1 and 3.0;
  its q0 is the breakup at the first parameter), Bf(L, q, q0, d) (L literal),
  q0(x[, md1, md2]).

* the daughters' momenta in the event frame: ``p1p``, ``p1e``, ``p1costheta``, ``p1phi``
  (|p|, E, p_z/|p|, atan2(p_y, p_x) of the first daughter) and ``p2...`` of the second;
* ``Flatte(m0, g_1..g_n, ma_1, mb_1, .., ma_n, mb_n)``, expanded as in the reference
  (src/CustomExpr.cu:499-511,844-870) into 1 / (m0^2 - m^2 + sum_i (g_i m0 / 2)
  (rho_i,im - i rho_i,re)) with the channel phase-space factors of the event mass.

``L``, ``d``, ``md1`` and ``md2`` are resolved to constants when the
plan is built (the GPU evaluates one lineshape per block, so ``L`` must be the same for
every wave at the node); ``md1``/``md2`` must be configured particle masses.

Serialized form (one token per node, prefix order): ``n:<value>`` number, ``I`` the
imaginary unit, ``v:m`` / ``v:q`` event variables, ``p:<k>`` parameter k, operators
``+ - * / ^ neg``, functions ``f:<name>`` (one argument), ``q0`` (3 arguments),
``bf:<L>`` (q, q0, d), ``bw`` (m, m0, g0), ``bwr:<L>`` (m, m0, g0, q, q0, d), ``rr`` /
``ri`` (ma, mb: the real / imaginary phase-space factor of the event mass), event
variables ``v:p1p`` .. ``v:p2phi``.
"""
from __future__ import annotations

import math
import re
from typing import Dict, List, Optional, Sequence, Tuple

from .config import ConfigError

_FUNCS = {"exp": "exp", "log": "log", "ln": "log", "sin": "sin", "cos": "cos", "sqrt": "sqrt",
          "csqrt": "sqrt", "abs": "abs", "Re": "re", "real": "re", "Im": "im", "imag": "im", "conj": "conj"}
_DAUGHTER_VARS = ("p1p", "p1e", "p1costheta", "p1phi", "p2p", "p2e", "p2costheta", "p2phi")

_TOKEN = re.compile(r"\s*(?:(1[ji])(?![A-Za-z0-9_])|(\d+\.?\d*(?:[eE][+-]?\d+)?|\.\d+(?:[eE][+-]?\d+)?)|"
                    r"([A-Za-z_][A-Za-z0-9_]*)|(\S))")


def tokenize(s: str) -> List[Tuple[str, object]]:
    out = []
    pos = 0
    while pos < len(s):
        m = _TOKEN.match(s, pos)
        if not m or m.end() == pos:
            break
        pos = m.end()
        imag, num, ident, ch = m.groups()
        if imag:
            out.append(("id", "1j"))
        elif num:
            out.append(("num", float(num)))
        elif ident:
            out.append(("id", ident))
        elif ch:
            if ch in "+-*/^(),":
                out.append(("op", ch))
            else:
                raise ConfigError("Custom expression: unexpected character %r" % ch)
    out.append(("end", None))
    return out



class PerWaveL(ConfigError):
    """The expression uses L and the waves at its node differ in L: the plan builder then
    parses it once per run of waves with one L (src/CustomExpr.cu:413, the reference's
    run-time per-wave L)."""

class Context:
    """What the expression may refer to at one node of one block."""

    def __init__(self, param_names: Sequence[str], L_values: Sequence[int], d: float,
                 md1: Optional[float], md2: Optional[float]):
        self.params = {n: k for k, n in enumerate(param_names)}
        self.L_values = sorted(set(L_values))
        self.d = d
        self.md1 = md1
        self.md2 = md2
        self.n_params = len(param_names)


class _Parser:
    def __init__(self, toks, ctx: Context):
        self.t = toks
        self.i = 0
        self.ctx = ctx

    def peek(self):
        return self.t[self.i]

    def take(self):
        tok = self.t[self.i]
        self.i += 1
        return tok

    def expect(self, op):
        tok = self.take()
        if tok != ("op", op):
            raise ConfigError("Custom expression: expected %r" % op)

    def parse(self):
        e = self.expr()
        if self.peek()[0] != "end":
            raise ConfigError("Custom expression: trailing tokens")
        return e

    def expr(self):
        left = self.term()
        while self.peek() in (("op", "+"), ("op", "-")):
            op = self.take()[1]
            left = (op, left, self.term())
        return left

    def term(self):
        left = self.factor()
        while self.peek() in (("op", "*"), ("op", "/")):
            op = self.take()[1]
            left = (op, left, self.factor())
        return left

    def factor(self):
        base = self.unary()
        if self.peek() == ("op", "^"):
            self.take()
            return ("^", base, self.factor())          # right associative
        return base

    def unary(self):
        if self.peek() == ("op", "-"):
            self.take()
            return ("neg", self.unary())
        if self.peek() == ("op", "+"):
            self.take()
            return self.unary()
        return self.primary()

    def args(self):
        self.expect("(")
        out = []
        if self.peek() == ("op", ")"):
            self.take()
            return out
        while True:
            out.append(self.expr())
            if self.peek() == ("op", ","):
                self.take()
                continue
            break
        self.expect(")")
        return out

    def q0_default(self):
        m0 = ("p", 0) if self.ctx.n_params > 0 else ("n", 1.0)
        return ("q0", m0, self.md(1), self.md(2))

    def md(self, k):
        v = self.ctx.md1 if k == 1 else self.ctx.md2
        if v is None:
            raise ConfigError("Custom expression: md%d is not a configured particle mass (not built)" % k)
        return ("n", v)

    @staticmethod
    def literal(node, what):
        if node[0] == "n":
            return node[1]
        if node[0] == "neg" and node[1][0] == "n":
            return -node[1][1]
        raise ConfigError("Custom expression: %s must be a literal" % what)

    def primary(self):
        kind, val = self.take()
        if kind == "num":
            return ("n", val)
        if kind == "op" and val == "(":
            e = self.expr()
            self.expect(")")
            return e
        if kind != "id":
            raise ConfigError("Custom expression: unexpected token %r" % (val,))
        name = val
        if name == "1j" or (name == "i" and name not in self.ctx.params):
            return ("I",)
        if name == "pi":
            return ("n", math.pi)
        if name == "m":
            return ("v", "m")
        if name == "q":
            return ("v", "q")
        if name == "q0" and self.peek() != ("op", "("):
            return self.q0_default()
        if name == "L":
            if len(self.ctx.L_values) != 1:
                raise PerWaveL("Custom expression uses L and the waves at its node have different L")
            return ("n", float(self.ctx.L_values[0]))
        if name == "d":
            return ("n", float(self.ctx.d))
        if name == "md1":
            return self.md(1)
        if name == "md2":
            return self.md(2)
        if name in _DAUGHTER_VARS:
            return ("v", name)
        if name in _FUNCS:
            a = self.args()
            if len(a) != 1:
                raise ConfigError("Custom expression: %s() takes one argument" % name)
            return ("f", _FUNCS[name], a[0])
        if name == "BW":
            a = self.args()
            if len(a) != 2:
                raise ConfigError("Custom expression: BW(m0, w0) needs 2 arguments")
            return ("bw", ("v", "m"), a[0], a[1])
        if name == "BWR":
            a = self.args()
            if not 2 <= len(a) <= 4:
                raise ConfigError("Custom expression: BWR(m0, g0[, L, d]) needs 2 to 4 arguments")
            L = int(self.literal(a[2], "BWR L")) if len(a) >= 3 else 1
            d = float(self.literal(a[3], "BWR d")) if len(a) >= 4 else 3.0
            return ("bwr", L, ("v", "m"), a[0], a[1], ("v", "q"), self.q0_default(), ("n", d))
        if name == "Bf":
            a = self.args()
            if len(a) != 4:
                raise ConfigError("Custom expression: Bf(L, q, q0, d) needs 4 arguments")
            L = int(self.literal(a[0], "Bf L"))
            return ("bf", L, a[1], a[2], a[3])
        if name == "q0":
            a = self.args()
            if len(a) == 1:
                return ("q0", a[0], self.md(1), self.md(2))
            if len(a) == 3:
                return ("q0", a[0], a[1], a[2])
            raise ConfigError("Custom expression: q0(x) or q0(x, md1, md2)")
        if name == "Flatte":
            a = self.args()
            if len(a) < 4 or (len(a) - 1) % 3 != 0:
                raise ConfigError("Custom expression: Flatte(m0, g..., (ma, mb)...) needs 1 + 3n arguments")
            n = (len(a) - 1) // 3
            m0, gs, ch = a[0], a[1:1 + n], a[1 + n:]
            m = ("v", "m")
            D = ("-", ("*", m0, m0), ("*", m, m))
            for i in range(n):
                gw = ("*", ("*", gs[i], m0), ("n", 0.5))
                rho = ("-", ("ri", ch[2 * i], ch[2 * i + 1]), ("*", ("I",), ("rr", ch[2 * i], ch[2 * i + 1])))
                D = ("+", D, ("*", gw, rho))
            return ("/", ("n", 1.0), D)
        if name in self.ctx.params:
            return ("p", self.ctx.params[name])
        raise ConfigError("Custom expression: unknown identifier '%s'" % name)


def parse(expression: str, ctx: Context):
    return _Parser(tokenize(expression), ctx).parse()


def serialize(node) -> List[str]:
    k = node[0]
    if k == "n":
        return ["n:%r" % float(node[1])]
    if k == "I":
        return ["I"]
    if k == "v":
        return ["v:" + node[1]]
    if k == "p":
        return ["p:%d" % node[1]]
    if k in ("+", "-", "*", "/", "^"):
        return [k] + serialize(node[1]) + serialize(node[2])
    if k == "neg":
        return ["neg"] + serialize(node[1])
    if k == "f":
        return ["f:" + node[1]] + serialize(node[2])
    if k == "q0":
        return ["q0"] + serialize(node[1]) + serialize(node[2]) + serialize(node[3])
    if k == "bf":
        return ["bf:%d" % node[1]] + serialize(node[2]) + serialize(node[3]) + serialize(node[4])
    if k == "bw":
        return ["bw"] + serialize(node[1]) + serialize(node[2]) + serialize(node[3])
    if k in ("rr", "ri"):
        return [k] + serialize(node[1]) + serialize(node[2])
    if k == "bwr":
        out = ["bwr:%d" % node[1]]
        for sub in node[2:]:
            out += serialize(sub)
        return out
    raise ValueError(node)


def uses_params(node) -> List[int]:
    out = set()

    def walk(n):
        if n[0] == "p":
            out.add(n[1])
        for x in n[1:]:
            if isinstance(x, tuple):
                walk(x)
    walk(node)
    return sorted(out)
