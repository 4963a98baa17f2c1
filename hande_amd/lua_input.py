"""Reader for the subset of HANDE's Lua input language used by FCIQMC decks.

HANDE drives calculations from Lua scripts such as::

    sys = hubbard_k { electrons = 4, lattice = { {4} }, ms = 0 }
    fciqmc { sys = sys, qmc = { tau = 0.01, rng_seed = 7, ... } }

(src/lua_hande.F90, src/lua_hande_systems.f90, src/lua_hande_calc.f90:1244-1503).
Lua is not available here, so this module parses the declarative subset these
decks use: assignments of calls with a single table argument, nested tables,
numbers, strings, booleans and references to earlier variables.  Comments
(``--`` and ``--[[ ... ]]``) are skipped.  Anything else raises ``ValueError``.
"""
import re

_TOKEN = re.compile(r"""
    (?P<ws>\s+)
  | (?P<lcomment>--\[\[.*?\]\])
  | (?P<comment>--[^\n]*)
  | (?P<number>[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?)
  | (?P<string>"[^"]*"|'[^']*')
  | (?P<name>[A-Za-z_][A-Za-z_0-9]*)
  | (?P<punct>[{}=,;()])
""", re.S | re.X)


class LuaCall:
    """A call ``func { ... }`` -- the table argument and the function name."""

    def __init__(self, func, args):
        self.func = func
        self.args = args

    def __repr__(self):
        return f"LuaCall({self.func!r}, {self.args!r})"


def _tokenize(text):
    pos = 0
    out = []
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m:
            raise ValueError(f"cannot parse input near: {text[pos:pos + 40]!r}")
        pos = m.end()
        kind = m.lastgroup
        if kind in ("ws", "comment", "lcomment"):
            continue
        out.append((kind, m.group(kind)))
    return out


class _Parser:
    def __init__(self, tokens):
        self.t = tokens
        self.i = 0
        self.env = {}
        self.calls = []

    def peek(self, k=0):
        j = self.i + k
        return self.t[j] if j < len(self.t) else (None, None)

    def take(self, value=None):
        tok = self.t[self.i]
        if value is not None and tok[1] != value:
            raise ValueError(f"expected {value!r}, found {tok[1]!r}")
        self.i += 1
        return tok

    def program(self):
        while self.i < len(self.t):
            kind, val = self.peek()
            if val == ";":
                self.take()
                continue
            if kind != "name":
                raise ValueError(f"unexpected token {val!r}")
            if self.peek(1)[1] == "=":
                name = self.take()[1]
                self.take("=")
                self.env[name] = self.expr()
            else:
                self.calls.append(self.expr())
        return self

    def expr(self):
        kind, val = self.peek()
        if val == "{":
            return self.table()
        if kind == "number":
            self.take()
            f = float(val)
            return int(f) if re.fullmatch(r"[-+]?\d+", val) else f
        if kind == "string":
            self.take()
            return val[1:-1]
        if kind == "name":
            self.take()
            if val in ("true", "false"):
                return val == "true"
            if val == "nil":
                return None
            nxt = self.peek()[1]
            if nxt == "{":
                return LuaCall(val, self.table())
            if nxt == "(":
                self.take("(")
                arg = self.table() if self.peek()[1] == "{" else {}
                self.take(")")
                return LuaCall(val, arg)
            if val in self.env:
                return self.env[val]
            raise ValueError(f"undefined name {val!r}")
        raise ValueError(f"unexpected token {val!r}")

    def table(self):
        self.take("{")
        keyed = {}
        seq = []
        while self.peek()[1] != "}":
            kind, val = self.peek()
            if kind == "name" and self.peek(1)[1] == "=":
                self.take()
                self.take("=")
                keyed[val] = self.expr()
            else:
                seq.append(self.expr())
            if self.peek()[1] in (",", ";"):
                self.take()
        self.take("}")
        if keyed and seq:
            keyed["_seq"] = seq
            return keyed
        return keyed if keyed else seq


def parse(text):
    """Parse a deck; returns (variables, list of top-level LuaCall)."""
    p = _Parser(_tokenize(text)).program()
    calls = list(p.calls)
    for v in p.env.values():
        if isinstance(v, LuaCall) and v not in calls:
            pass
    return p.env, calls
