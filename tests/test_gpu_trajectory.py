"""GPU: the time loop over a WRF run directory, HIP dycore vs CPU oracle.

A seeded 28x24x12 case (8 species, PD advection, hill) is written as a run
directory with three wrfbdy records 54 s apart whose edge values drift with
nonzero tendencies, then run for six steps with snapshots every 36 s
(src/wrf_jax/trajectory.py:3933-4700 drives the same sequence: chunks that
end at record and output times, dtbc restarted at 0 and the next record
loaded at every record time).  The same loop over the oracle is the
reference: every registry field after the run within the one-step bar, and
every snapshot equal to the oracle's.
"""
import numpy as np
import pytest

from parity import REL_L2_STEP, STEP_DIAGNOSTICS, compare

pytestmark = pytest.mark.gpu


def test_trajectory_with_boundary_records(gpu, tmp_path):
    from test_real_case import OracleDomain, _records
    from wrfdc.case_files import export_case
    from wrfdc.cases import make_case
    from wrfdc.data_io import load_case
    from wrfdc.fields import FIELD_SPECS
    from wrfdc.real_case import dycore_config
    from wrfdc.trajectory import run_trajectory

    print("[trajectory] shape=(28, 24, 12), 3 records, 6 steps", flush=True)
    case = make_case(28, 24, 12, nmoist=6, nscalar=2, hydrometeors=True, hill_height=500.0)
    export_case(tmp_path, case.config, case.arrays, records=_records(case, 3), interval_seconds=54, run_seconds=108)
    cd = load_case(tmp_path)
    names = tuple(n for n, _ in FIELD_SPECS)
    got = run_trajectory(cd, snapshot_interval_seconds=36, device=gpu, final_fields=names)
    dom = OracleDomain(dycore_config(cd))
    ref = run_trajectory(cd, snapshot_interval_seconds=36, domain=dom, final_fields=names)
    dom.close()
    assert got.metrics["boundary_refreshes"] == ref.metrics["boundary_refreshes"] == 2
    assert [s.elapsed_seconds for s in got.snapshots] == [36, 72, 108]
    bad = compare(got.final_arrays, ref.final_arrays, tol=REL_L2_STEP, skip=STEP_DIAGNOSTICS)
    assert not bad, f"fields above rel-L2 {REL_L2_STEP} after the trajectory: {bad}"
    for a, b in zip(got.snapshots, ref.snapshots):
        for f in ("u", "v", "t", "qvapor"):
            x, y = getattr(a, f), getattr(b, f)
            err = float(np.linalg.norm((x - y).ravel()) / max(np.linalg.norm(y.ravel()), 1e-30))
            assert err <= REL_L2_STEP, (a.elapsed_seconds, f, err)
