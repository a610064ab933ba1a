"""``namelist.input`` -> case configuration.

Interface of src/wrf_jax/case_config.py (``CaseConfig``, ``parse_namelist``,
``load_case_config``).  The namelist is read by a lexer over the whole text
and a small state machine over its tokens, which covers the Fortran namelist
grammar WRF's ``namelist.input`` uses: ``&group`` ... ``/`` (or ``&end``)
blocks, ``key = v1, v2, ...`` assignments whose values may run over several
lines, ``!`` comments, logicals (``.true.``, ``.f.``, ``t`` ...), quoted
strings (a doubled quote stands for one), integers, reals with ``e`` or ``d``
exponents and ``r*value`` repeats.  Names are case-insensitive and stored in
lower case.  A key with one value maps to it, with several to their list;
domain-indexed quantities are read for domain 1 (index 0).
"""

from __future__ import annotations

import re
from dataclasses import dataclass
from pathlib import Path
from typing import Any, Iterator


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


_LEX = re.compile(
    r"""
      (?P<space>\s+)
    | (?P<comment>![^\n]*)
    | (?P<open>&(?P<gname>[A-Za-z_]\w*))
    | (?P<close>/)
    | (?P<key>[A-Za-z_][\w%]*(?:\(\s*\d+\s*\))?)\s*=
    | (?P<str>'(?:[^']|'')*'|"(?:[^"]|"")*")
    | (?P<comma>,)
    | (?P<atom>[^\s,/!&'"=]+)
    """,
    re.VERBOSE,
)

_LOGICAL = {".true.": True, ".t.": True, "t": True, "true": True, ".false.": False, ".f.": False, "f": False, "false": False}
_INT = re.compile(r"[+-]?\d+$")
_REAL = re.compile(r"[+-]?(\d+\.?\d*|\.\d+)([ed][+-]?\d+)?$", re.IGNORECASE)


def _lex(text: str, where: str) -> Iterator[tuple[str, str]]:
    pos = 0
    while pos < len(text):
        m = _LEX.match(text, pos)
        if m is None:
            line = text.count("\n", 0, pos) + 1
            raise ValueError(f"{where}:{line}: cannot read namelist text {text[pos:pos + 20]!r}")
        pos = m.end()
        kind = m.lastgroup
        if kind in ("space", "comment"):
            continue
        if kind in ("open", "gname"):
            yield "open", m.group("gname").lower()
        elif kind == "key":
            yield "key", re.sub(r"\s+", "", m.group("key")).lower()
        else:
            yield kind, m.group(kind)


def _scalar(atom: str) -> Any:
    low = atom.lower()
    if low in _LOGICAL:
        return _LOGICAL[low]
    if _INT.match(atom):
        return int(atom)
    if _REAL.match(atom):
        return float(low.replace("d", "e"))
    return atom


def _atom_values(atom: str) -> list:
    """One unquoted item; ``r*v`` stands for r copies of v."""
    count, star, rest = atom.partition("*")
    if star and count.isdigit() and rest:
        return [_scalar(rest)] * int(count)
    return [_scalar(atom)]


def _unquote(tok: str) -> str:
    q = tok[0]
    return tok[1:-1].replace(q + q, q)


def parse_namelist(path: str | Path) -> dict:
    """Groups (lower-case names) -> {key (lower case): value or list}."""
    where = str(path)
    groups: dict[str, dict[str, Any]] = {}
    current: dict[str, Any] | None = None
    values: list | None = None
    for kind, tok in _lex(Path(path).read_text(encoding="ascii"), where):
        if kind == "open":
            if tok == "end":  # '&end' closes a group in some namelist dialects
                current = values = None
            else:
                current = groups.setdefault(tok, {})
                values = None
        elif kind == "close":
            current = values = None
        elif current is None:
            continue  # text between groups is ignored, as Fortran does
        elif kind == "key":
            values = current[tok] = []
        elif values is None:
            raise ValueError(f"{where}: value {tok!r} before any key in a namelist group")
        elif kind == "str":
            values.append(_unquote(tok))
        elif kind == "atom":
            values.extend(_atom_values(tok))
    return {g: {k: (v[0] if len(v) == 1 else v) for k, v in kv.items()} for g, kv in groups.items()}


def _domain1(namelist: dict, group: str, key: str, default: Any = None) -> Any:
    """The domain-1 value of ``group.key`` (first entry of a list)."""
    v = namelist.get(group, {}).get(key, default)
    return v[0] if isinstance(v, list) and v else v


def _time_stamp(namelist: dict, which: str) -> str:
    parts = [int(_domain1(namelist, "time_control", f"{which}_{unit}", 0)) for unit in
             ("year", "month", "day", "hour", "minute", "second")]
    return "{:04d}-{:02d}-{:02d}_{:02d}:{:02d}:{:02d}".format(*parts)


# history output interval keys in the order WRF honours them, with seconds
# per unit; the plain key counts minutes
_HISTORY_KEYS = (("history_interval_s", 1), ("history_interval_m", 60), ("history_interval_h", 3600),
                 ("history_interval", 60))


def _history_interval(namelist: dict) -> int:
    for key, seconds in _HISTORY_KEYS:
        v = _domain1(namelist, "time_control", key)
        if v is not None:
            return int(v) * seconds
    return 60 * 60


def load_case_config(path: str | Path) -> CaseConfig:
    nl = parse_namelist(path)
    run = {unit: int(_domain1(nl, "time_control", f"run_{unit}", 0)) for unit in ("days", "hours", "minutes", "seconds")}
    run_seconds = ((run["days"] * 24 + run["hours"]) * 60 + run["minutes"]) * 60 + run["seconds"]

    def need(group: str, key: str):
        v = _domain1(nl, group, key)
        if v is None:
            raise KeyError(f"{path}: &{group} has no {key}")
        return v

    return CaseConfig(
        namelist=nl,
        run_seconds=run_seconds,
        time_step=float(need("domains", "time_step")),
        boundary_interval_seconds=int(_domain1(nl, "time_control", "interval_seconds", run_seconds)),
        history_interval_seconds=_history_interval(nl),
        max_dom=int(_domain1(nl, "domains", "max_dom", 1)),
        e_we=int(need("domains", "e_we")),
        e_sn=int(need("domains", "e_sn")),
        e_vert=int(need("domains", "e_vert")),
        dx=float(need("domains", "dx")),
        dy=float(need("domains", "dy")),
        start_time=_time_stamp(nl, "start"),
        end_time=_time_stamp(nl, "end"),
    )


__all__ = ["CaseConfig", "load_case_config", "parse_namelist"]
