"""WRF run directories: namelist, wrfinput / wrfbdy NetCDF, the time loop.

Covers the input side of the em_real driver (src/wrf_jax/case_config.py,
src/wrf_jax/data_io.py, src/wrf_jax/runtime.py:4312-4431 field assembly) and
its loop (src/wrf_jax/trajectory.py: boundary-record refresh at
``interval_seconds``, dtbc reset, snapshots, history comparison).  The
reference ships no case files, so cases are written by ``case_files`` from
the seeded synthetic workloads and read back; the time loop is driven over
the CPU oracle here and over the HIP dycore in the GPU tests.
"""
import dataclasses

import numpy as np
import pytest

from wrfdc.case_config import load_case_config, parse_namelist
from wrfdc.case_files import export_case
from wrfdc.cases import make_case
from wrfdc.data_io import (assert_state_finite, load_case, load_wrfbdy, load_wrfbdy_fields, load_wrfinput,
                           load_wrfinput_fields)
from wrfdc.fields import FIELD_SPECS, Dims
from wrfdc.real_case import boundary_record, dycore_config, registry_arrays
from wrfdc.step_em import TimeState, from_registry
from wrfdc.trajectory import (compare_snapshots_to_wrfout, next_boundary_reset_step, run_trajectory,
                              write_snapshots_netcdf)

NX, NY, NZ = 16, 14, 8


def _case():
    return make_case(name="rc", nx=NX, ny=NY, nz=NZ, config_index=0, nmoist=6, nscalar=2, hydrometeors=True,
                     hill_height=500.0)


def _records(case, n, seed=7):
    """n boundary records: record 0 from the case, later ones drift by up
    to 5 % with nonzero tendencies."""
    rng = np.random.default_rng(seed)
    names = [k for k, kind in FIELD_SPECS if kind in "XYxy" and "_b" in k and "_bt" not in k]
    recs = []
    cur = {k: case.arrays[k].copy() for k in names}
    for r in range(n):
        rec = {}
        for k in names:
            t = k.replace("_b", "_bt", 1)
            rec[k] = cur[k].copy()
            rec[t] = (cur[k] * rng.uniform(-0.05, 0.05, cur[k].shape) / 600.0).astype(np.float32)
        recs.append(rec)
        cur = {k: (rec[k] + rec[k.replace("_b", "_bt", 1)] * np.float32(54.0)).astype(np.float32) for k in names}
    return recs


class OracleDomain:
    """The CPU oracle behind the Domain methods the time loop calls."""

    def __init__(self, cfg):
        from oracle import Oracle

        self.cfg, self.Oracle, self.o, self.calls = cfg, Oracle, None, []

    def upload(self, arrays):
        if self.o is None:
            self.o = self.Oracle(self.cfg, dict(arrays))
        else:
            for k, v in arrays.items():
                self.o.arrays[k][...] = v
        self.calls.append(("upload", sorted(arrays)[:1]))

    def step(self, n, dtbc0):
        self.calls.append(("step", n, dtbc0))
        self.o.step(n, dtbc0)

    def download(self, names):
        return {n: self.o.arrays[n].copy() for n in names}

    def close(self):
        self.o.close()


def test_namelist_parsing(tmp_path):
    p = tmp_path / "namelist.input"
    p.write_text("""! leading comment
 &time_control
 run_days = 0,
 run_hours = 1, ! one hour
 run_minutes = 30,
 start_year = 2024, 2024,
 start_month = 06,
 start_day = 15,
 start_hour = 12,
 end_year = 2024,
 end_month = 06,
 end_day = 15,
 end_hour = 13,
 end_minute = 30,
 interval_seconds = 10800
 history_interval_m = 20, 60,
 history_outname = 'wrfout_d<domain>_!x, y',
 /
 &domains
 time_step = 18,
 max_dom = 2,
 e_we = 101, 202,
 e_sn = 91,
 e_vert = 46,
 dx = 3000.d0, 1000.,
 dy = 3.0e3,
 eta_levels = 1.0, 0.99,
  0.98, 0.97,
 /
 &dynamics
 non_hydrostatic = .true.,
 use_theta_m = .FALSE.,
 /
""")
    g = parse_namelist(p)
    assert g["time_control"]["history_outname"] == "wrfout_d<domain>_!x, y"
    assert g["domains"]["eta_levels"] == [1.0, 0.99, 0.98, 0.97]
    assert g["domains"]["e_we"] == [101, 202] and g["domains"]["dx"] == [3000.0, 1000.0]
    assert g["dynamics"]["non_hydrostatic"] is True and g["dynamics"]["use_theta_m"] is False
    c = load_case_config(p)
    assert c.run_seconds == 5400 and c.time_step == 18.0 and c.boundary_interval_seconds == 10800
    assert c.history_interval_seconds == 1200
    assert (c.max_dom, c.e_we, c.e_sn, c.e_vert, c.dx, c.dy) == (2, 101, 91, 46, 3000.0, 3000.0)
    assert c.start_time == "2024-06-15_12:00:00" and c.end_time == "2024-06-15_13:30:00"


def test_namelist_grammar(tmp_path):
    """Repeat counts, doubled quotes, '&end' terminators, text between
    groups, comments after values and keys spread over several lines."""
    p = tmp_path / "namelist.input"
    p.write_text("""&domains
 e_we = 3*101,
 e_sn = 91 e_vert = 46 ! two assignments on one line
 dx = 1.5D3, dy = 1500
 title = 'it''s "quoted"' ,
&end
stray text outside any group
&physics
 mp_physics = 0, 0,
   0 /
""")
    g = parse_namelist(p)
    assert g["domains"]["e_we"] == [101, 101, 101]
    assert g["domains"]["e_sn"] == 91 and g["domains"]["e_vert"] == 46
    assert g["domains"]["dx"] == 1500.0 and g["domains"]["dy"] == 1500
    assert g["domains"]["title"] == 'it\'s "quoted"'
    assert g["physics"]["mp_physics"] == [0, 0, 0]
    bad = tmp_path / "bad.input"
    bad.write_text("&domains\n 7, 8\n/\n")
    with pytest.raises(ValueError):
        parse_namelist(bad)


def test_case_files_round_trip(tmp_path):
    case = _case()
    export_case(tmp_path, case.config, case.arrays, interval_seconds=54)
    cd = load_case(tmp_path)
    cfg = dycore_config(cd)
    want = dataclasses.asdict(case.config)
    got = dataclasses.asdict(cfg)
    for k, v in want.items():
        if k == "options":
            continue
        if k in ("cf1", "cf2", "cf3", "cfn", "cfn1", "ptop"):  # stored as FP32 scalars in the file
            assert got[k] == float(np.float32(v)), k
        else:
            assert got[k] == v, k
    reg = registry_arrays(cd, cfg)
    d = Dims(nx=NX, ny=NY, nz=NZ, nmoist=6, nscalar=2)
    ext3 = {"u_2": (NX + 1, NZ, NY), "v_2": (NX, NZ, NY + 1), "w_2": (NX, NZ + 1, NY), "t_2": (NX, NZ, NY),
            "ph_2": (NX, NZ + 1, NY), "phb": (NX, NZ + 1, NY), "p": (NX, NZ, NY), "pb": (NX, NZ, NY), "al": (NX, NZ, NY),
            "alb": (NX, NZ, NY)}
    for k, (ni, nk, nj) in ext3.items():
        a = from_registry(k, case.arrays[k], d)[:ni, :nk, :nj]
        b = from_registry(k, reg[k], d)[:ni, :nk, :nj]
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), k
        full = from_registry(k, reg[k], d)
        # horizontal padding repeats the last column / row (runtime.py:3907-3935)
        assert np.array_equal(full[ni:, :nk, :nj], np.repeat(full[ni - 1 : ni, :nk, :nj], full.shape[0] - ni, axis=0)), k
    for k in ("mub", "ht", "msftx", "msfty", "msfux", "msfvx", "msfvx_inv", "cosa"):
        ni = NX + 1 if k in ("msfux",) else NX
        nj = NY + 1 if k in ("msfvx", "msfvx_inv") else NY
        a = from_registry(k, case.arrays[k], d)[:ni, :nj]
        b = from_registry(k, reg[k], d)[:ni, :nj]
        assert np.array_equal(a, b), k
    for k in ("znw", "dnw", "rdnw", "dn", "rdn", "fnm", "fnp", "c1h", "c2h", "c1f", "c2f", "c3h", "c3f"):
        n = NZ + 1 if k in ("znw", "c1f", "c2f", "c3f") else NZ
        assert np.array_equal(reg[k][:n], case.arrays[k][:n]), k
    for s in range(6):
        a = case.arrays["moist"][s, :NY, :NZ, :NX]
        assert np.array_equal(reg["moist"][s, :NY, :NZ, :NX], a), s
    for k in ("u_bxs", "u_bye", "t_bxe", "ph_bys", "qv_bxs", "mu_bxs", "mu_bye"):
        kind = dict(FIELD_SPECS)[k]
        sl = {"X": np.s_[:, :NZ, :NY], "Y": np.s_[:, :NZ, :NX], "x": np.s_[:, :NY], "y": np.s_[:, :NX]}[kind]
        assert np.array_equal(reg[k][sl], case.arrays[k][sl]), k
    grid, state = load_wrfinput(tmp_path / "wrfinput_d01", cd.config)
    assert (grid.west_east, grid.south_north, grid.bottom_top, grid.west_east_stag) == (NX, NY, NZ, NX + 1)
    assert state.u.shape == (NZ, NY, NX + 1)
    assert_state_finite(state)
    fields = load_wrfinput_fields(tmp_path / "wrfinput_d01")
    assert fields["u"].shape == (NX + 1, NZ, NY) and fields["mub"].shape == (NX, NY)


def test_dycore_from_files_matches_synthetic_case(tmp_path):
    """Two steps of the oracle from the file-loaded state equal two steps from
    the synthetic arrays on every physical point (the padding points the
    loader fills by edge replication are never read)."""
    from oracle import Oracle

    case = _case()
    export_case(tmp_path, case.config, case.arrays, interval_seconds=54)
    cd = load_case(tmp_path)
    cfg = dycore_config(cd)
    full = {k: v.copy() for k, v in case.arrays.items()}
    full.update(registry_arrays(cd, cfg))
    o1, o2 = Oracle(case.config, case.arrays), Oracle(cfg, full)
    o1.step(2, 0.0)
    o2.step(2, 0.0)
    ext = {"u_2": (NY, NZ, NX + 1), "v_2": (NY + 1, NZ, NX), "w_2": (NY, NZ + 1, NX), "t_2": (NY, NZ, NX),
           "ph_2": (NY, NZ + 1, NX), "p": (NY, NZ, NX), "al": (NY, NZ, NX)}
    for k, (nj, nk, ni) in ext.items():
        assert np.array_equal(o1.arrays[k][:nj, :nk, :ni], o2.arrays[k][:nj, :nk, :ni]), k
    assert np.array_equal(o1.arrays["mu_2"][:NY, :NX], o2.arrays["mu_2"][:NY, :NX])
    for k in ("moist", "scalar"):
        assert np.array_equal(o1.arrays[k][:, :NY, :NZ, :NX], o2.arrays[k][:, :NY, :NZ, :NX]), k
    o1.close()
    o2.close()


def test_wrfbdy_records(tmp_path):
    case = _case()
    recs = _records(case, 3)
    export_case(tmp_path, case.config, case.arrays, records=recs, interval_seconds=54)
    f = load_wrfbdy(tmp_path / "wrfbdy_d01")
    assert f.times == ("2000-01-01_00:00:00", "2000-01-01_00:00:54", "2000-01-01_00:01:48")
    assert "U_BXS" in f.names and f.arrays[f.names.index("U_BXS")].shape == (3, 5, NZ, NY)
    r1 = load_wrfbdy_fields(tmp_path / "wrfbdy_d01", time_index=1)
    assert r1["u_bxs"].shape == (NY, NZ, 5) and r1["mu_btxs"].shape == (NY, 1, 5)
    d = Dims(nx=NX, ny=NY, nz=NZ, nmoist=6, nscalar=2)
    reg = boundary_record(tmp_path / "wrfbdy_d01", 2, d)
    assert np.array_equal(reg["t_btye"][:, :NZ, :NX], recs[2]["t_btye"][:, :NZ, :NX])
    assert np.array_equal(reg["qv_bxs"][:, :NZ, :NY], recs[2]["qv_bxs"][:, :NZ, :NY])
    with pytest.raises(IndexError):
        load_wrfbdy_fields(tmp_path / "wrfbdy_d01", time_index=3)


def test_input_errors(tmp_path):
    case = _case()
    arrays = {k: v for k, v in case.arrays.items() if k != "al"}
    export_case(tmp_path / "noal", case.config, arrays)
    with pytest.raises(KeyError, match="AL"):
        registry_arrays(load_case(tmp_path / "noal"), case.config)
    export_case(tmp_path / "diff", case.config, case.arrays)
    nl = tmp_path / "diff" / "namelist.input"
    nl.write_text(nl.read_text().replace("diff_opt = 2", "diff_opt = 1"))
    with pytest.raises(NotImplementedError, match="diff_opt"):
        dycore_config(load_case(tmp_path / "diff"))
    nl.write_text(nl.read_text().replace("diff_opt = 1", "diff_opt = 2").replace("moist_adv_opt = 1", "moist_adv_opt = 2"))
    with pytest.raises(NotImplementedError, match="moist_adv_opt"):
        dycore_config(load_case(tmp_path / "diff"))
    nl.write_text(nl.read_text().replace("moist_adv_opt = 2", "moist_adv_opt = 1"))
    for line, key in ((" spec_exp = 0.33,", "spec_exp"), (" have_bcs_moist = .true.,", "have_bcs_moist")):
        nl.write_text(nl.read_text().replace(" specified = .true.,", " specified = .true.,\n" + line))
        with pytest.raises(NotImplementedError, match=key):
            dycore_config(load_case(tmp_path / "diff"))
        nl.write_text(nl.read_text().replace(line + "\n", ""))
    dycore_config(load_case(tmp_path / "diff"))
    h5 = tmp_path / "h5"
    h5.mkdir()
    (h5 / "wrfinput_d01").write_bytes(b"\x89HDF\r\n\x1a\n" + bytes(64))
    with pytest.raises(NotImplementedError, match="HDF5"):
        load_wrfinput_fields(h5 / "wrfinput_d01")


def test_next_boundary_reset_step():
    # 54 s records at dt 18: every third step
    assert next_boundary_reset_step(1, 10, 18, 54) == 3
    assert next_boundary_reset_step(3, 10, 18, 54) == 3
    assert next_boundary_reset_step(4, 10, 18, 54) == 6
    assert next_boundary_reset_step(10, 10, 18, 54) is None
    assert next_boundary_reset_step(1, 10, 18, 0) is None
    assert next_boundary_reset_step(1, 10, 18, 10) is None


def test_trajectory_over_oracle(tmp_path):
    """Six steps with a record every three: the loop runs 2 + 1 + 1 + 2
    steps (snapshots every 36 s, records every 54 s), restarts dtbc at 0
    after steps 3 and 6, loads record 1 after step 3 and record 2 after step
    6 (as the reference does at every record time the file holds), and ends
    with the state of a hand-driven oracle run."""
    from oracle import Oracle

    case = _case()
    recs = _records(case, 3)
    export_case(tmp_path, case.config, case.arrays, records=recs, interval_seconds=54, run_seconds=108)
    cd = load_case(tmp_path)
    cfg = dycore_config(cd)
    dom = OracleDomain(cfg)
    res = run_trajectory(cd, snapshot_interval_seconds=36, domain=dom, final_fields=("u_2", "t_2", "moist", "ph_2"))
    steps = [c[1:] for c in dom.calls if c[0] == "step"]
    assert steps == [(2, 0.0), (1, 36.0), (1, 0.0), (2, 18.0)]
    assert res.metrics["boundary_refreshes"] == 2 and res.metrics["steps"] == 6
    assert [s.elapsed_seconds for s in res.snapshots] == [36, 72, 108]
    assert res.snapshots[-1].timestamp == "2000-01-01_00:01:48"
    assert res.time.itimestep == 6 and float(res.time.dtbc) == 0.0
    # the same run by hand
    o = Oracle(cfg, registry_arrays(cd, cfg))
    o.step(3, 0.0)
    for k, v in boundary_record(tmp_path / "wrfbdy_d01", 1, Dims(nx=NX, ny=NY, nz=NZ, nmoist=6, nscalar=2)).items():
        o.arrays[k][...] = v
    o.step(3, 0.0)
    for k in ("u_2", "t_2", "moist", "ph_2"):
        assert np.array_equal(res.final_arrays[k], o.arrays[k]), k
    snap = res.snapshots[-1]
    assert np.array_equal(snap.t, o.arrays["t_2"][:NY, :NZ, :NX].transpose(1, 0, 2))
    assert np.array_equal(snap.qvapor, o.arrays["moist"][0, :NY, :NZ, :NX].transpose(1, 0, 2))
    o.close()
    dom.close()
    # history file round trip and comparison
    out = tmp_path / "wrfout_d01"
    write_snapshots_netcdf(res.snapshots, out)
    cmp = compare_snapshots_to_wrfout(res.snapshots, out)
    assert len(cmp) == 12 and all(c.rmse == 0.0 and c.nrmse_percent == 0.0 and c.finite for c in cmp)
    moved = [dataclasses.replace(s, t=s.t + np.float32(1.0)) for s in res.snapshots]
    cmp = {(c.elapsed_seconds, c.field): c for c in compare_snapshots_to_wrfout(tuple(moved), out)}
    c = cmp[(108, "T")]
    # T + 1 K rounds in FP32 (|T| up to ~40 K): errors of 1 K within 4e-6
    assert abs(c.rmse - 1.0) < 4e-6 and abs(c.max_abs_error - 1.0) < 4e-6
    ref = res.snapshots[-1].t.astype(np.float64)
    assert abs(c.nrmse_percent / (100.0 / np.sqrt(np.mean(ref * ref))) - 1.0) < 4e-6
    assert cmp[(108, "U")].rmse == 0.0


def test_time_state_advance_resets_dtbc():
    t = TimeState(itimestep=np.int32(0), elapsed_seconds=np.float32(0.0), dtbc=np.float32(0.0))
    seen = []
    for _ in range(4):
        t = t.advance(18.0, 36)
        seen.append(float(t.dtbc))
    assert seen == [18.0, 0.0, 18.0, 0.0] and int(t.itimestep) == 4
