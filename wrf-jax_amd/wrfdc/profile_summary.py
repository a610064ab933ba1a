"""Text summaries of rocprofv3 result databases (SQLite) for ``profiles/``.

Five reports:

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

``traffic_vs_model FETCH_DB WRITE_DB TRACE_DB``
    the measured bytes per launch beside the algorithmic bytes of
    ``bytes_model.KERNELS`` for the 400x400x60 workload, with the mean launch
    time of a kernel trace of the same command.

``issue COUNTER_DB TRACE_DB``
    instruction issue per kernel from one ``--pmc`` run of ``SQ_WAVES
    SQ_INSTS_VALU SQ_INSTS_VMEM_RD SQ_INSTS_SALU SQ_INSTS_LDS
    SQ_ACTIVE_INST_VALU GRBM_GUI_ACTIVE``: VALU and vector-memory
    instructions per wave, the engine clock over the dispatch
    (GRBM_GUI_ACTIVE cycles of one XCD / the trace's mean duration) and the share of the
    dispatch the SIMDs spend issuing VALU work: SQ_INSTS_VALU x 4 cycles per
    wave64 instruction, and SQ_ACTIVE_INST_VALU (quad-cycles), over 1024
    SIMDs x GRBM_GUI_ACTIVE.

``mfma COUNTER_DB TRACE_DB``
    matrix-core use of the kernels that issue MFMA instructions, from one
    ``--pmc`` run of ``SQ_WAVES SQ_INSTS_MFMA SQ_VALU_MFMA_BUSY_CYCLES
    SQ_INSTS_VALU_TRANS_F32 GRBM_GUI_ACTIVE``.

Usage: ``python -m wrfdc.profile_summary kernels DB``,
``python -m wrfdc.profile_summary traffic FETCH_DB WRITE_DB``,
``python -m wrfdc.profile_summary traffic_vs_model FETCH_DB WRITE_DB TRACE_DB``,
``python -m wrfdc.profile_summary issue COUNTER_DB TRACE_DB`` or
``python -m wrfdc.profile_summary mfma COUNTER_DB TRACE_DB``.
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


def traffic_vs_model(fetch_db: str, write_db: str, trace_db: str, workload: str = "c1") -> str:
    """Per kernel of bytes_model.KERNELS: mean launch time (kernel trace),
    measured HBM bytes per launch (the two counter runs, calibrated as in
    ``traffic``), the algorithmic bytes of the model, their ratio and the
    bandwidth each implies."""
    from .bytes_model import KERNELS
    from .cases import CONFIGS

    c = CONFIGS[workload]
    nmoist, nscalar = c["nmoist"], c["nscalar"]
    npts = c["nx"] * c["ny"] * c["nz"]
    meas = {}
    for line in traffic(fetch_db, write_db).splitlines():
        parts = [p.strip() for p in line.strip().strip("|").split("|")]
        if len(parts) == 5 and parts[3].replace(".", "").isdigit():
            meas[parts[0]] = float(parts[3])
    cur = sqlite3.connect(trace_db).cursor()
    dur = {_short(n): avg for n, calls, tot, avg in cur.execute("select name, total_calls, total_duration, average from top_kernels")}
    out = [f"Workload {workload}: {c['nx']}x{c['ny']}x{c['nz']}, {nmoist} moist + {nscalar} scalar species "
           f"(algorithmic bytes: 4 B x N_mass per distinct 3-D field pass, bytes_model.KERNELS)", "",
           "| kernel | mean us | traffic MB | algorithmic MB | traffic / alg | GB/s (traffic) | GB/s (alg) |",
           "|---|---:|---:|---:|---:|---:|---:|"]
    rows = []
    def match(name, names):
        # a model entry without template arguments stands for the one
        # instance of that kernel that has no entry of its own
        if name in names:
            return name
        if "<" in name:
            return None
        return next((k for k in names if k.startswith(name + "<") and k not in KERNELS), None)

    for name, km in KERNELS.items():
        key, dk = match(name, meas), match(name, dur)
        if key is None or dk is None:
            continue
        alg = 4.0 * npts * km.passes(nmoist, nscalar)
        us = dur[dk]
        rows.append((us, name, meas[key], alg))
    for us, name, tr, alg in sorted(rows, reverse=True):
        out.append(f"| {name} | {us:.1f} | {tr / 1e6:.0f} | {alg / 1e6:.0f} | {tr / alg:.2f} | {tr / (us * 1e3):.0f} | "
                   f"{alg / (us * 1e3):.0f} |")
    return "\n".join(out)


N_SIMD = 1024  # MI355X: 256 compute units x 4 SIMDs
N_XCD = 8  # GRBM_GUI_ACTIVE comes summed over the 8 XCDs, each counting the dispatch's cycles


def issue(counter_db: str, trace_db: str) -> str:
    cur = sqlite3.connect(counter_db).cursor()
    per = collections.defaultdict(lambda: collections.defaultdict(float))
    name = {}
    for kname, cname, v, disp in cur.execute("select kernel_name, counter_name, value, dispatch_id from counters_collection"):
        d = per[disp]
        d[cname] += v
        name[disp] = _short(kname)
    tcur = sqlite3.connect(trace_db).cursor()
    dur = {_short(n): avg for n, avg in tcur.execute("select name, average from top_kernels")}
    agg = collections.defaultdict(lambda: collections.defaultdict(float))
    cnt = collections.Counter()
    for disp, d in per.items():
        for c, v in d.items():
            agg[name[disp]][c] += v
        cnt[name[disp]] += 1
    out = ["| kernel | dispatches | mean us | clock GHz | VALU / wave | VMEM rd / wave | VALU issue % | VALU active % |",
           "|---|---:|---:|---:|---:|---:|---:|---:|"]
    rows = []
    for k, a in agg.items():
        n = cnt[k]
        waves = a["SQ_WAVES"] / n
        cyc = a["GRBM_GUI_ACTIVE"] / n / N_XCD
        if waves <= 0 or cyc <= 0 or k not in dur:
            continue
        us = dur[k]
        valu = a["SQ_INSTS_VALU"] / n
        rows.append((us * n, k, n, us, cyc / (us * 1e3), valu / waves, a["SQ_INSTS_VMEM_RD"] / n / waves,
                     100.0 * valu * 4.0 / (N_SIMD * cyc), 100.0 * a["SQ_ACTIVE_INST_VALU"] / n * 4.0 / (N_SIMD * cyc)))
    for _, k, n, us, ghz, vw, mw, vi, va in sorted(rows, reverse=True):
        out.append(f"| {k} | {n} | {us:.1f} | {ghz:.2f} | {vw:.0f} | {mw:.0f} | {vi:.0f} | {va:.0f} |")
    return "\n".join(out)


def mfma(counter_db: str, trace_db: str) -> str:
    """Matrix-core use of the kernels that issue MFMA instructions, from one
    counter pass of SQ_WAVES, SQ_INSTS_MFMA, SQ_VALU_MFMA_BUSY_CYCLES,
    SQ_INSTS_VALU_TRANS_F32 and GRBM_GUI_ACTIVE: MFMA and transcendental
    instructions per wave and the share of SIMD cycles the MFMA unit is busy
    (busy cycles summed over the SIMDs / (1024 SIMDs x one XCD's cycles))."""
    cur = sqlite3.connect(counter_db).cursor()
    per = collections.defaultdict(lambda: collections.defaultdict(float))
    name = {}
    for kname, cname, v, disp in cur.execute("select kernel_name, counter_name, value, dispatch_id from counters_collection"):
        per[disp][cname] += v
        name[disp] = _short(kname)
    tcur = sqlite3.connect(trace_db).cursor()
    dur = {_short(n): avg for n, avg in tcur.execute("select name, average from top_kernels")}
    agg = collections.defaultdict(lambda: collections.defaultdict(float))
    cnt = collections.Counter()
    for disp, d in per.items():
        for c, v in d.items():
            agg[name[disp]][c] += v
        cnt[name[disp]] += 1
    out = ["| kernel | dispatches | mean us | MFMA / wave | transcendental / wave | MFMA busy % |",
           "|---|---:|---:|---:|---:|---:|"]
    rows = []
    for k, a in agg.items():
        n = cnt[k]
        waves = a["SQ_WAVES"] / n
        cyc = a["GRBM_GUI_ACTIVE"] / n / N_XCD
        if a["SQ_INSTS_MFMA"] <= 0 or waves <= 0 or cyc <= 0:
            continue
        us = dur.get(k, float("nan"))
        rows.append((us * n, k, n, us, a["SQ_INSTS_MFMA"] / n / waves, a["SQ_INSTS_VALU_TRANS_F32"] / n / waves,
                     100.0 * a["SQ_VALU_MFMA_BUSY_CYCLES"] / n / (N_SIMD * cyc)))
    for _, k, n, us, mw, tw, busy in sorted(rows, reverse=True):
        out.append(f"| {k} | {n} | {us:.1f} | {mw:.0f} | {tw:.0f} | {busy:.1f} |")
    return "\n".join(out)


def main(argv):
    if len(argv) >= 2 and argv[0] == "kernels":
        print(kernels(argv[1]))
    elif len(argv) >= 3 and argv[0] == "traffic":
        print(traffic(argv[1], argv[2]))
    elif len(argv) >= 4 and argv[0] == "traffic_vs_model":
        print(traffic_vs_model(argv[1], argv[2], argv[3]))
    elif len(argv) >= 3 and argv[0] == "issue":
        print(issue(argv[1], argv[2]))
    elif len(argv) >= 3 and argv[0] == "mfma":
        print(mfma(argv[1], argv[2]))
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
