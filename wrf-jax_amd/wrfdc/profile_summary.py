"""Text summaries of rocprofv3 result databases (SQLite) for ``profiles/``.

Two reports:

``kernels DB``
    per-kernel calls, total and mean duration from a ``--kernel-trace --stats``
    run (the ``top_kernels`` view).

``traffic FETCH_DB WRITE_DB``
    HBM bytes per dispatch from two ``--pmc`` runs (``FETCH_SIZE``, then
    ``WRITE_SIZE``) of ``bench.py --traffic-calibration``.  The last dispatch
    of each run is a 256 MiB device copy of known traffic (256 MiB read,
    256 MiB written); each counter is scaled by known / measured for that copy.
    Rows: ``| kernel | fetch MiB | write MiB | traffic bytes | dispatches |``,
    averaged over the kernel's dispatches (the format ``bytes_model.
    measured_traffic`` reads).

Usage: ``python -m wrfdc.profile_summary kernels DB`` or
``python -m wrfdc.profile_summary traffic FETCH_DB WRITE_DB``.
"""

from __future__ import annotations

import collections
import sqlite3
import sys

CAL_BYTES = 256 << 20  # bench.py --traffic-calibration: 64 Mi float32


def _short(name: str) -> str:
    return name.split("(")[0].replace("void ", "").strip()


def kernels(db: str) -> str:
    cur = sqlite3.connect(db).cursor()
    rows = list(cur.execute("select name, total_calls, total_duration, average from top_kernels"))
    total = sum(r[2] for r in rows)
    # the top_kernels view reports durations in microseconds
    out = [f"Total kernel time {total / 1e3:.2f} ms", "", "| kernel | calls | total us | avg us | % |", "|---|---:|---:|---:|---:|"]
    for name, calls, dur, avg in rows:
        out.append(f"| {_short(name)} | {calls} | {dur:.0f} | {avg:.1f} | {100.0 * dur / total:.2f} |")
    return "\n".join(out)


def _per_dispatch(db: str, counter: str):
    cur = sqlite3.connect(db).cursor()
    val = collections.defaultdict(float)
    name = {}
    for kname, cname, v, disp in cur.execute("select kernel_name, counter_name, value, dispatch_id from counters_collection"):
        if cname != counter:
            continue
        val[disp] += v
        name[disp] = _short(kname)
    return val, name


def traffic(fetch_db: str, write_db: str) -> str:
    fv, fn = _per_dispatch(fetch_db, "FETCH_SIZE")
    wv, wn = _per_dispatch(write_db, "WRITE_SIZE")
    cal_f = max(fv)
    cal_w = max(wv)
    f_scale = CAL_BYTES / (fv[cal_f] * 1024.0)
    w_scale = CAL_BYTES / (wv[cal_w] * 1024.0)
    per = collections.defaultdict(lambda: [0.0, 0.0, 0])
    for d, v in fv.items():
        if d == cal_f:
            continue
        per[fn[d]][0] += v * 1024.0 * f_scale
        per[fn[d]][2] += 1
    wcount = collections.Counter()
    for d, v in wv.items():
        if d == cal_w:
            continue
        per[wn[d]][1] += v * 1024.0 * w_scale
        wcount[wn[d]] += 1
    out = [
        f"Calibration (last dispatch, {fn[cal_f]}, {CAL_BYTES >> 20} MiB copy): FETCH_SIZE read "
        f"{fv[cal_f] / 1024:.1f} MiB -> factor {f_scale:.3f}; WRITE_SIZE read {wv[cal_w] / 1024:.1f} MiB -> factor {w_scale:.3f}.",
        "",
        "| kernel | fetch MiB | write MiB | traffic bytes | dispatches |",
        "|---|---:|---:|---:|---:|",
    ]
    rows = []
    for k, (fb, wb, n) in per.items():
        nf = max(n, 1)
        nw = max(wcount[k], 1)
        fm, wm = fb / nf, wb / nw
        rows.append((fm + wm, k, fm, wm, n))
    for tot, k, fm, wm, n in sorted(rows, reverse=True):
        out.append(f"| {k} | {fm / 2**20:.1f} | {wm / 2**20:.1f} | {tot:.0f} | {n} |")
    return "\n".join(out)


def main(argv):
    if len(argv) >= 2 and argv[0] == "kernels":
        print(kernels(argv[1]))
    elif len(argv) >= 3 and argv[0] == "traffic":
        print(traffic(argv[1], argv[2]))
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
