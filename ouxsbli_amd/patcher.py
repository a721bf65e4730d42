"""Text patching of a case's ``config.fypp`` macros and ``mod_globals.f90``
parameters -- the same contract as the reference's ``ouxsbli.patcher``
(ouxsbli/patcher.py:7-139): a changed name is rewritten in place, trailing
``!`` comments survive, Python values are spelled the Fortran/fypp way, and a
name absent from a fypp file is appended as a new ``#:set`` line (never to a
plain Fortran file).
"""
import re

__all__ = ["patch"]


def _fortran_double(v: float) -> str:
    """Python float -> Fortran double literal: 1600.0 -> '1600.0d0', 1e-5 -> '1d-05'."""
    text = f"{v:.10g}"
    if "e" in text or "E" in text:
        return text.replace("E", "d").replace("e", "d")
    if "." not in text:
        text += ".0"
    return text + "d0"


# name kept for callers of the reference module
_to_fortran_double = _fortran_double


def _split_comment(line: str):
    cut = line.find("!")
    return (line, "") if cut < 0 else (line[:cut], line[cut:])


def _is_param_line(line: str, name: str) -> bool:
    """``[type][, attrs], parameter :: name ...`` (case-insensitive, code part only)."""
    code, _ = _split_comment(line)
    return re.search(r"parameter\s*::\s*" + re.escape(name) + r"\b", code, re.IGNORECASE) is not None


def _is_macro_line(line: str, name: str) -> bool:
    """``#:set name = ...`` (case-insensitive)."""
    return re.search(r"^\s*#:\s*set\s+" + re.escape(name) + r"\b", line, re.IGNORECASE) is not None


def _param_value(value) -> str:
    if isinstance(value, bool):
        return ".true." if value else ".false."
    if isinstance(value, int):
        return str(value)
    if isinstance(value, float):
        return _fortran_double(value)
    return str(value)


def _macro_value(value) -> str:
    if isinstance(value, bool):
        return "True" if value else "False"
    if isinstance(value, int):
        return str(value)
    if isinstance(value, float):
        return _fortran_double(value)
    if isinstance(value, str):
        v = value.strip()
        quoted = len(v) >= 2 and v[0] == v[-1] and v[0] in "'\""
        return v if quoted else f"'{v}'"
    return str(value)


def _rewrite_param(line: str, name: str, value) -> str:
    code, comment = _split_comment(line)
    new = re.sub(r"(parameter\s*::\s*" + re.escape(name) + r"\s*=\s*).*$",
                 lambda m: m.group(1) + _param_value(value), code, flags=re.IGNORECASE)
    return new + comment


def _rewrite_macro(line: str, name: str, value) -> str:
    code, comment = _split_comment(line)
    new = re.sub(r"(#:\s*set\s+" + re.escape(name) + r"\s*=\s*)(.+?)\s*$",
                 lambda m: m.group(1) + _macro_value(value), code, flags=re.IGNORECASE)
    if new == code and "=" in code:  # empty right-hand side
        new = code.split("=")[0] + "= " + _macro_value(value)
    return new + comment


def patch(src_text: str, changes: dict) -> str:
    """Return *src_text* with every name in *changes* set to its new value."""
    lines = src_text.splitlines()
    for n, line in enumerate(lines):
        for name, value in changes.items():
            if _is_macro_line(line, name):
                lines[n] = _rewrite_macro(line, name, value)
                break
            if _is_param_line(line, name):
                lines[n] = _rewrite_param(line, name, value)
                break
    if any("#:set" in l for l in lines):
        for name, value in changes.items():
            if not any(_is_macro_line(l, name) or _is_param_line(l, name) for l in lines):
                lines.append(_rewrite_macro(f"#:set {name} = ", name, value))
    return "\n".join(lines) + "\n"
