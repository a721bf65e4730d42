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


def _py_expr(expr: str) -> str:
    e = expr.strip()
    e = re.sub(r"(\d+\.?\d*|\.\d+)[dD]([+-]?\d+)", lambda m: m.group(1) + "e" + m.group(2), e)  # 1.d0 -> 1.e0
    e = re.sub(r"(\d|\.)_(sp|dp|\d+)\b", r"\1", e)  # kind suffixes 0.4_sp
    e = re.sub(r"\.true\.", "True", e, flags=re.I)
    e = re.sub(r"\.false\.", "False", e, flags=re.I)
    e = re.sub(r"\bkind\s*=\s*\w+", "", e)
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


def read_parameters(text: str) -> dict:
    """``... parameter :: name = expr`` lines -> {name (lower case): value}.
    Lines that cannot be evaluated (derived types, unknown functions) are skipped."""
    env: dict = {}
    for raw in text.splitlines():
        code = raw.split("!")[0]
        m = re.search(r"parameter\s*::\s*(\w+)\s*=\s*(.+)$", code, re.I)
        if not m:
            continue
        name, expr = m.group(1).lower(), m.group(2)
        if "dim3" in expr.lower():
            continue
        try:
            py = _py_expr(expr).lower()
            env[name] = eval(py, {"__builtins__": {}, **_FUNCS}, dict(env))
        except Exception:
            continue
    return env
