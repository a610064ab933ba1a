"""GPU tests of the batched ensemble (one captured graph for the RK3 step of
uneven members, csrc/wrfdc.hip wrfdc_ensemble_*).

* replaying the graph advances every member exactly as stepping it alone
  does (bit-identical fields), also with time-varying lateral forcing, where
  each replay must see the current boundary time;
* every member matches the FP32 oracle after the replays (relative L2 1e-5),
  also the smallest members of the 32-member batched workload;
* the autotuner returns a valid layout and tile shape.
"""

import numpy as np
import pytest

from parity import REL_L2_STEP, STEP_DIAGNOSTICS, compare

pytestmark = pytest.mark.gpu

SPECS = [  # nx, ny, nz, dx, dt
    (24, 20, 10, 2000.0, 12.0),
    (30, 26, 12, 3000.0, 18.0),
    (36, 30, 14, 5000.0, 30.0),
    (28, 32, 11, 1500.0, 9.0),
]


def _cases():
    from wrfdc.cases import make_case

    out = []
    for m, (nx, ny, nz, dx, dt) in enumerate(SPECS):
        c = make_case(nx, ny, nz, config_index=2, member=m, dx=dx, dt=dt, time_step_sound=0, nmoist=3, nscalar=1,
                      hydrometeors=True, hill_height=400.0)
        rng = np.random.default_rng(100 + m)
        # time-varying lateral forcing: the relaxation zones see bdy + dtbc * tendency
        for name in ("u_btxs", "u_btxe", "v_btys", "v_btye", "t_btxs", "t_btye", "mu_btxs", "qv_btys"):
            c.arrays[name][:] = (1e-3 * rng.standard_normal(c.arrays[name].shape)).astype(np.float32)
        out.append(c)
    return out


def _domains(cases, gpu):
    from wrfdc.device import Domain

    doms = []
    for c in cases:
        d = Domain(c.config, gpu)
        d.upload(c.arrays)
        doms.append(d)
    return doms


def test_graph_replay_matches_members_stepped_alone(gpu):
    from oracle import Oracle
    from wrfdc.ensemble import Ensemble

    print("[ensemble] 4 uneven members, 2 streams, 3 replays vs eager steps and the oracle", flush=True)
    cases = _cases()
    A = _domains(cases, gpu)
    B = _domains(cases, gpu)
    ens = Ensemble(A, branch_of=[0, 1, 0, 1], nbranches=2)
    assert ens.nodes > 0
    dtbc0 = 5.0
    ens.step(3, dtbc0)
    ens.synchronize()
    for m, (a, b, c) in enumerate(zip(A, B, cases)):
        b.step(3, dtbc0)
        ga, gb = a.download(), b.download()
        for name in ga:
            assert np.array_equal(ga[name].view(np.uint32), gb[name].view(np.uint32)), (m, name)
        o = Oracle(c.config, c.arrays)
        o.step(3, dtbc0)
        bad = compare(ga, o.arrays, tol=REL_L2_STEP, skip=STEP_DIAGNOSTICS)
        o.close()
        assert not bad, (m, bad)
    ens.close()
    for d in A + B:
        d.close()


def test_workload_members_match_oracle(gpu):
    """The three smallest members of the 32-member batched workload
    (104^2..116^2 columns, 47..58 levels, 6 moist species, PD on, their own
    dx and dt) advanced two steps by one captured ensemble graph on two
    streams: every member within relative L2 1e-5 of the FP32 oracle."""
    from oracle import Oracle
    from wrfdc.device import Domain
    from wrfdc.ensemble import Ensemble, c2_members, member_case

    specs = c2_members()
    pick = sorted(range(len(specs)), key=lambda q: specs[q].nx * specs[q].ny * specs[q].nz)[:3]
    print(f"[ensemble] workload members {pick}, 2 streams, 2 replays vs the oracle", flush=True)
    cases = [member_case(specs[q], q) for q in pick]
    doms = []
    for c in cases:
        d = Domain(c.config, gpu)
        d.upload(c.arrays)
        doms.append(d)
    ens = Ensemble(doms, branch_of=[0, 1, 0], nbranches=2)
    ens.step(2, 0.0)
    ens.synchronize()
    ens.close()
    for q, d, c in zip(pick, doms, cases):
        got = d.download()
        d.close()
        o = Oracle(c.config, c.arrays)
        o.step(2, 0.0)
        bad = compare(got, o.arrays, tol=REL_L2_STEP, skip=STEP_DIAGNOSTICS)
        o.close()
        assert not bad, (q, bad)


def test_autotune_returns_a_layout(gpu):
    from wrfdc.ensemble import Ensemble, autotune

    print("[ensemble] autotune over 1, 2, 4 streams and face_levels 0, 2, 16", flush=True)
    doms = _domains(_cases(), gpu)
    nb, assign, table, tiles = autotune(doms, candidates=(1, 2, 4), reps=2, face_levels=(0, 2, 16),
                                        log=lambda s: print("[ensemble]", s, flush=True))
    assert nb in (1, 2, 4) and len(assign) == len(doms) and all(0 <= b < nb for b in assign)
    assert len(table) == 3 and all(t > 0 for _, t in table)
    assert [lg for lg, _ in tiles] == [0, 2, 16] and all(t > 0 for _, t in tiles)
    ens = Ensemble(doms, assign, nb)
    ens.step(1)
    ens.synchronize()
    ens.close()
    for d in doms:
        assert d.checksums()[3] == 0
        d.close()
