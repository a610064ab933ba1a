"""CPU tests of the measurement helpers: the traffic table reader, the
calibrated rocprofv3 summary and bench.py's single-GPU guard."""

import os
import sqlite3
import subprocess
import sys

from wrfdc import bytes_model, profile_summary

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_step_model_matches_documented_workload():
    # 400x400x60, 8 species, 4 acoustic substeps: 28.2 GB per RK3 step
    assert abs(bytes_model.step_bytes(400 * 400 * 60, 8) / 1e9 - 28.19) < 0.01


def test_measured_traffic_reads_rows_with_template_arguments(tmp_path):
    p = tmp_path / "traffic_x.md"
    # the traffic-vs-model table comes first in the profile and has more
    # columns: its rows must not be taken for the raw counter rows
    p.write_text("| kernel | mean us | traffic MB | algorithmic MB | traffic / alg | GB/s (traffic) | GB/s (alg) |\n"
                 "|---|---:|---:|---:|---:|---:|---:|\n"
                 "| k_acoustic_col | 437.6 | 1791 | 1382 | 1.30 | 4093 | 3159 |\n\n"
                 "| kernel | fetch MiB | write MiB | traffic bytes | dispatches |\n"
                 "|---|---:|---:|---:|---:|\n"
                 "| k_acoustic_col<8> | 1.0 | 2.0 | 3.5e+09 | 9 |\n"
                 "| k_uv_substep | 1.0 | 2.0 | 5.0e+08 | 7 |\n")
    assert bytes_model.measured_traffic("k_acoustic_col", "x", str(p)) == 3.5e9
    assert bytes_model.measured_traffic("k_uv_substep", "x", str(p)) == 5.0e8
    assert bytes_model.measured_traffic("k_missing", "x", str(p)) is None


def _counter_db(path, counter, rows):
    con = sqlite3.connect(path)
    con.execute("create table counters_collection (kernel_name text, counter_name text, value real, dispatch_id int)")
    con.executemany("insert into counters_collection values (?, ?, ?, ?)", rows)
    con.commit()
    con.close()


def test_traffic_summary_scales_by_the_calibration_copy(tmp_path):
    f, w = str(tmp_path / "f.db"), str(tmp_path / "w.db")
    kib = 1024.0
    # kernel A: two dispatches, each split over two XCD entries; the last
    # dispatch is the 256 MiB calibration copy, read at half its bytes
    _counter_db(f, "FETCH_SIZE", [("kA(int)", "FETCH_SIZE", 50 * kib, 1), ("kA(int)", "FETCH_SIZE", 50 * kib, 1),
                                  ("kA(int)", "FETCH_SIZE", 100 * kib, 2), ("copy", "FETCH_SIZE", 128 * kib, 3)])
    _counter_db(w, "WRITE_SIZE", [("kA(int)", "WRITE_SIZE", 30 * kib, 1), ("kA(int)", "WRITE_SIZE", 30 * kib, 2),
                                  ("copy", "WRITE_SIZE", 256 * kib, 3)])
    text = profile_summary.traffic(f, w)
    assert "factor 2.000" in text and "factor 1.000" in text
    row = [ln for ln in text.splitlines() if ln.startswith("| kA |")][0]
    cells = [c.strip() for c in row.strip("|").split("|")]
    assert float(cells[1]) == 200.0  # 100 MiB per dispatch x 2
    assert float(cells[2]) == 30.0
    assert float(cells[3]) == 230.0 * 2**20
    assert cells[4] == "2"


def test_mfma_report_busy_share(tmp_path):
    c, t = str(tmp_path / "c.db"), str(tmp_path / "t.db")
    con = sqlite3.connect(c)
    con.execute("create table counters_collection (kernel_name text, counter_name text, value real, dispatch_id int)")
    rows = []
    for disp in (1, 2):  # two dispatches, GRBM summed over the 8 XCDs
        rows += [("kM(x)", "SQ_WAVES", 1000.0, disp), ("kM(x)", "SQ_INSTS_MFMA", 64000.0, disp),
                 ("kM(x)", "SQ_INSTS_VALU_TRANS_F32", 266000.0, disp), ("kM(x)", "GRBM_GUI_ACTIVE", 8 * 2000.0, disp),
                 ("kM(x)", "SQ_VALU_MFMA_BUSY_CYCLES", 0.25 * 1024 * 2000.0, disp),
                 ("kV(x)", "SQ_WAVES", 10.0, 3), ("kV(x)", "SQ_INSTS_MFMA", 0.0, 3), ("kV(x)", "GRBM_GUI_ACTIVE", 80.0, 3)]
    con.executemany("insert into counters_collection values (?, ?, ?, ?)", rows)
    con.commit()
    con.close()
    con = sqlite3.connect(t)
    con.execute("create table top_kernels (name text, total_calls int, total_duration real, average real)")
    con.execute("insert into top_kernels values ('kM(x)', 2, 2000.0, 1000.0)")
    con.commit()
    con.close()
    text = profile_summary.mfma(c, t)
    rows = [ln for ln in text.splitlines() if ln.startswith("| kM") or ln.startswith("| kV")]
    assert len(rows) == 1  # the kernel without MFMA instructions is left out
    cells = [x.strip() for x in rows[0].strip("|").split("|")]
    assert cells[:6] == ["kM", "2", "1000.0", "64", "266", "25.0"]


def test_bench_refuses_more_than_one_gpu():
    res = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert res.returncode == 2
    assert "one GPU" in res.stderr
