"""Reads a case's compile-time configuration: the ``#:set`` macros of
``config.fypp`` (fypp evaluates them as Python expressions) and the
``parameter ::`` declarations of ``mod_globals.f90`` (Fortran constant
expressions, evaluated in declaration order)."""
from __future__ import annotations

import math
import re

_FUNCS = {
    "sqrt": math.sqrt, "dsqrt": math.sqrt, "sin": math.sin, "dsin": math.sin, "cos": math.cos, "dcos": math.cos,
    "tan": math.tan, "dtan": math.tan, "atan": math.atan, "datan": math.atan, "acos": math.acos, "dacos": math.acos,
    "asin": math.asin, "dasin": math.asin, "dexp": math.exp, "dlog": math.log,
    "exp": math.exp, "log": math.log, "abs": abs, "dabs": abs, "dble": float, "real": lambda x, *a, **k: float(x),
    "int": int, "nint": round, "min": min, "max": max, "mod": math.fmod, "kind": lambda x: 8,
}


class _FInt(int):
    """A Fortran default integer inside a constant expression: ``/`` between
    two integers truncates toward zero (Fortran 2008, 7.1.5.2.1), and integer
    results of + - * ** and unary minus stay integers."""

    def _wrap(self, v):
        return _FInt(v) if isinstance(v, int) and not isinstance(v, bool) else v

    def __truediv__(self, o):
        if isinstance(o, int) and not isinstance(o, bool):
            if o == 0:
                raise ZeroDivisionError("integer division by zero")
            q = abs(int(self)) // abs(int(o))
            return _FInt(q if (int(self) >= 0) == (int(o) >= 0) else -q)
        return int(self) / o

    def __rtruediv__(self, o):
        if isinstance(o, int) and not isinstance(o, bool):
            return _FInt(o).__truediv__(self)
        return o / int(self)

    def __add__(self, o): return self._wrap(int(self) + o)
    def __radd__(self, o): return self._wrap(o + int(self))
    def __sub__(self, o): return self._wrap(int(self) - o)
    def __rsub__(self, o): return self._wrap(o - int(self))
    def __mul__(self, o): return self._wrap(int(self) * o)
    def __rmul__(self, o): return self._wrap(o * int(self))
    def __neg__(self): return _FInt(-int(self))
    def __pos__(self): return self

    def __pow__(self, o):
        if isinstance(o, int) and o < 0:  # integer ** negative integer: 1/(i**|n|), truncated
            return _FInt(1).__truediv__(_FInt(int(self) ** (-o)))
        return self._wrap(int(self) ** o)

    def __rpow__(self, o): return self._wrap(o ** int(self))


# Tokens of a constant expression: names, real literals (1.0, .5, 1e3, 1.e-3
# after the d -> e rewrite) and integer literals; only the last are wrapped.
_TOKEN = re.compile(r"[A-Za-z_]\w*|(\d+\.\d*|\.\d+)([eE][+-]?\d+)?|\d+[eE][+-]?\d+|(\d+)")


def _py_expr(expr: str) -> str:
    e = expr.strip()
    e = re.sub(r"(\d|\.)_(sp|dp|\d+)\b", r"\1", e)  # kind suffixes 0.4_sp
    e = re.sub(r"(\d+\.?\d*|\.\d+)[dD]([+-]?\d+)", lambda m: m.group(1) + "e" + m.group(2), e)  # 1.d0 -> 1.e0
    e = re.sub(r"\.true\.", "True", e, flags=re.I)
    e = re.sub(r"\.false\.", "False", e, flags=re.I)
    e = re.sub(r"\bkind\s*=\s*\w+", "", e)
    e = _TOKEN.sub(lambda m: f"_I({m.group(3)})" if m.group(3) else m.group(0), e)
    return e


def read_fypp_macros(text: str) -> dict:
    """``#:set NAME = expr`` lines -> {NAME: value} (fypp semantics: Python expressions)."""
    env: dict = {}
    for line in text.splitlines():
        m = re.match(r"^\s*#:\s*set\s+(\w+)\s*=\s*(.+?)\s*$", line)
        if not m:
            continue
        name, expr = m.group(1), m.group(2)
        try:
            env[name] = eval(expr, {"__builtins__": {}}, dict(env))
        except Exception:
            env[name] = expr.strip().strip("'\"")
    return env


# Parameters Case reads from mod_globals.f90: a line declaring one of these
# that cannot be evaluated is an error, not a silent skip.
REQUIRED = ("nx", "ny", "nz", "dt", "lx", "ly", "lz", "nt", "np")


def read_parameters(text: str, required=REQUIRED) -> dict:
    """``... parameter :: name = expr`` lines -> {name (lower case): value},
    evaluated in declaration order with Fortran semantics: integer literals
    and integer-declared names are integers, ``/`` between integers
    truncates toward zero, and each value takes its declared type
    (``integer`` truncates, ``real``/``double precision`` converts).  Lines
    that cannot be evaluated (derived types, unknown functions) are skipped,
    unless they declare one of ``required``, which raises ValueError."""
    env: dict = {}
    for raw in text.splitlines():
        code = raw.split("!")[0]
        m = re.search(r"^(.*?)parameter\s*::\s*(\w+)\s*=\s*(.+)$", code, re.I)
        if not m:
            continue
        decl, name, expr = m.group(1).lower(), m.group(2).lower(), m.group(3)
        if "dim3" in expr.lower():
            continue
        try:
            py = _py_expr(expr).lower().replace("_i(", "_I(")
            v = eval(py, {"__builtins__": {}, "_I": _FInt, **_FUNCS}, dict(env))
            if re.match(r"\s*integer", decl):
                v = _FInt(int(v))
            elif re.match(r"\s*(real|double)", decl) and not isinstance(v, bool):
                v = float(v)
            env[name] = v
        except Exception as e:
            if name in required:
                raise ValueError(f"cannot evaluate parameter {name!r} = {expr.strip()!r}: {e}") from e
            continue
    return {k: (int(v) if isinstance(v, _FInt) else v) for k, v in env.items()}
