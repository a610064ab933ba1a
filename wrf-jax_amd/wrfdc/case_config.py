"""``namelist.input`` -> case configuration.

Mirrors src/wrf_jax/case_config.py:1-190 (``CaseConfig``, ``parse_namelist``,
``load_case_config``): the subset of Fortran namelist syntax that WRF's
``namelist.input`` uses (``&group`` ... ``/`` blocks, ``key = v1, v2, ...``
assignments that may continue over lines, ``!`` comments outside quotes,
``.true.``/``.false.``, quoted strings, integers and reals with ``d``/``e``
exponents).  A key with one value maps to that value, with several to the
list; every domain-indexed quantity is read at index 0 (domain 1).
"""

from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path
from typing import Any


@dataclass(frozen=True)
class CaseConfig:
    """Runtime configuration parsed from ``namelist.input``."""

    namelist: dict
    run_seconds: int
    time_step: float
    boundary_interval_seconds: int
    history_interval_seconds: int
    max_dom: int
    e_we: int
    e_sn: int
    e_vert: int
    dx: float
    dy: float
    start_time: str
    end_time: str


def _code_part(line: str) -> str:
    """The line up to its first ``!`` that is not inside a quoted string."""
    quote = None
    for pos, ch in enumerate(line):
        if quote:
            if ch == quote:
                quote = None
        elif ch in "'\"":
            quote = ch
        elif ch == "!":
            return line[:pos]
    return line


def _tokens(text: str) -> list[str]:
    """Comma-separated values of one assignment (commas inside quotes kept)."""
    out, cur, quote = [], [], None
    for ch in text:
        if quote:
            cur.append(ch)
            if ch == quote:
                quote = None
        elif ch in "'\"":
            quote = ch
            cur.append(ch)
        elif ch == ",":
            tok = "".join(cur).strip()
            if tok:
                out.append(tok)
            cur = []
        else:
            cur.append(ch)
    tok = "".join(cur).strip()
    if tok:
        out.append(tok)
    return out


def _value(tok: str) -> Any:
    low = tok.lower()
    if low in (".true.", "true", ".t."):
        return True
    if low in (".false.", "false", ".f."):
        return False
    if len(tok) >= 2 and tok[0] == tok[-1] and tok[0] in "'\"":
        return tok[1:-1]
    try:
        if any(c in low for c in ".ed"):
            return float(low.replace("d", "e"))
        return int(tok)
    except ValueError:
        return tok


def parse_namelist(path: str | Path) -> dict:
    """Groups (lower-case names) -> {key (lower case): value or list}."""
    groups: dict[str, dict[str, Any]] = {}
    group: str | None = None
    key: str | None = None
    parts: list[str] = []

    def close() -> None:
        nonlocal key, parts
        if group is not None and key is not None:
            vals = [_value(t) for t in _tokens(" ".join(parts))]
            groups[group][key] = vals[0] if len(vals) == 1 else vals
        key, parts = None, []

    for raw in Path(path).read_text(encoding="ascii").splitlines():
        line = _code_part(raw).strip()
        if not line:
            continue
        if line.startswith("&"):
            close()
            group = line[1:].strip().lower()
            groups[group] = {}
        elif line.startswith("/"):
            close()
            group = None
        elif group is None:
            continue
        elif "=" in line:
            close()
            lhs, rhs = line.split("=", 1)
            key, parts = lhs.strip().lower(), [rhs.strip()]
        elif key is not None:
            parts.append(line)
    close()
    return groups


def _get(groups: dict, group: str, key: str, default: Any = None, index: int = 0) -> Any:
    """Value of ``group.key`` at domain ``index`` (the last entry when the list
    is shorter), or ``default``."""
    v = groups.get(group, {}).get(key, default)
    if isinstance(v, list):
        return v[index] if index < len(v) else v[-1]
    return v


def _stamp(groups: dict, prefix: str) -> str:
    f = lambda name, default=None: int(_get(groups, "time_control", f"{prefix}_{name}", default))  # noqa: E731
    return f"{f('year'):04d}-{f('month'):02d}-{f('day'):02d}_{f('hour', 0):02d}:{f('minute', 0):02d}:{f('second', 0):02d}"


def _history_seconds(groups: dict) -> int:
    tc = groups.get("time_control", {})
    for key, scale in (("history_interval_s", 1), ("history_interval_m", 60), ("history_interval_h", 3600)):
        if tc.get(key) is not None:
            return int(_get(groups, "time_control", key)) * scale
    return int(_get(groups, "time_control", "history_interval", 60)) * 60


def load_case_config(path: str | Path) -> CaseConfig:
    g = parse_namelist(path)
    tc = lambda key: int(_get(g, "time_control", key, 0))  # noqa: E731
    run_seconds = tc("run_days") * 86400 + tc("run_hours") * 3600 + tc("run_minutes") * 60 + tc("run_seconds")
    return CaseConfig(
        namelist=g,
        run_seconds=run_seconds,
        time_step=float(_get(g, "domains", "time_step")),
        boundary_interval_seconds=int(_get(g, "time_control", "interval_seconds", run_seconds)),
        history_interval_seconds=_history_seconds(g),
        max_dom=int(_get(g, "domains", "max_dom", 1)),
        e_we=int(_get(g, "domains", "e_we")),
        e_sn=int(_get(g, "domains", "e_sn")),
        e_vert=int(_get(g, "domains", "e_vert")),
        dx=float(_get(g, "domains", "dx")),
        dy=float(_get(g, "domains", "dy")),
        start_time=_stamp(g, "start"),
        end_time=_stamp(g, "end"),
    )


__all__ = ["CaseConfig", "load_case_config", "parse_namelist"]
