"""Mutation check of the oracle's pins.

Each case below plants one plausible slip in a copy of the oracle's C
sources (a flipped sign, a dropped term, a wrong coefficient or index),
builds that copy in a temporary folder, and runs the oracle pins
(test_oracle_closed_forms.py and test_oracle_pinning.py) against it.  Every
slip must make at least one pin fail, and the unmodified copy must pass
them all: otherwise a misreading of that kind in the restatement, which the
bit-exact GPU parity tests would inherit, could go unnoticed.

CPU only; about a minute on 8 cores.
"""

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
SRC = os.path.join(REPO, "oracle", "src")
PINS = [os.path.join(HERE, "test_oracle_closed_forms.py"), os.path.join(HERE, "test_oracle_pinning.py")]
CFLAGS = ["-O1", "-fPIC", "-shared", "-std=c11", "-ffp-contract=off", "-w"]
SOURCES = ["o_bc.c", "o_advect.c", "o_bigstep.c", "o_small.c", "o_diff.c", "o_step.c"]

# (id, file, text, replacement, occurrence: 0-based index of the match to change)
MUTATIONS = [
    # the ten slips of the review that asked for these pins
    ("coef_w_bu_minus_bl", "o_small.c", "* (bu + bl);", "* (bu - bl);", 0),
    ("coef_w_offcentre_sign", "o_small.c", "real cof = (real)0.5 * dts * g * ((real)1 + epssm);",
     "real cof = (real)0.5 * dts * g * ((real)1 - epssm);", 0),
    ("smag_cs_not_squared", "o_diff.c", "real km = c_s * c_s * mlen * mlen * sd;", "real km = c_s * mlen * mlen * sd;", 0),
    ("advect_u_x_divergence_sign", "o_advect.c",
     "for (int i = 2; i <= ide - 1; i++) A3(tend, i, k, j) -= A2(msfux, i, j) * rdx * (fqx[i + 1] - fqx[i]);",
     "for (int i = 2; i <= ide - 1; i++) A3(tend, i, k, j) += A2(msfux, i, j) * rdx * (fqx[i + 1] - fqx[i]);", 0),
    ("advect_u_vertical_upwind_direction", "o_advect.c",
     "A3(u, i, k, j), A3(u, i, k + 1, j), -vel);", "A3(u, i, k, j), A3(u, i, k + 1, j), vel);", 0),
    ("relax_laplacian_sign_rows", "o_bc.c", "x->fcx[bd + 1] * f0 - x->gcx[bd + 1]", "x->fcx[bd + 1] * f0 + x->gcx[bd + 1]", 0),
    ("relax_laplacian_sign_columns", "o_bc.c", "x->fcx[bd + 1] * f0 - x->gcx[bd + 1]", "x->fcx[bd + 1] * f0 + x->gcx[bd + 1]", 1),
    ("coriolis_u_sign", "o_bigstep.c", "val = val + (A2(msfux, i, j) / A2(msfuy, i, j)) * (real)0.5 * (A2(f, i, j)",
     "val = val - (A2(msfux, i, j) / A2(msfuy, i, j)) * (real)0.5 * (A2(f, i, j)", 0),
    ("hdiff_u_sign", "o_diff.c", "A3(tend, i, k, j) = A3(tend, i, k, j) + inc;", "A3(tend, i, k, j) = A3(tend, i, k, j) - inc;", 0),
    ("stress_sign", "o_diff.c", "A3(tt, i, k, j) = -A3(rho, i, k, j) * A3(xk, i, k, j) * A3(defor, i, k, j);",
     "A3(tt, i, k, j) = A3(rho, i, k, j) * A3(xk, i, k, j) * A3(defor, i, k, j);", 0),
    ("sumflux_drops_ww_lin", "o_small.c", "if (fin) acc = acc / ns + A3(ww_lin, i, k, j);", "if (fin) acc = acc / ns;", 0),
    ("advance_uv_emdiv_sign", "o_small.c", "real damping = (real)-1 * emdiv * grid_length", "real damping = (real)1 * emdiv * grid_length", 0),
    # further slips of the same kinds
    ("advect_v_vertical_upwind_direction", "o_advect.c",
     "A3(v, i, k, j), A3(v, i, k + 1, j), -vel);", "A3(v, i, k, j), A3(v, i, k + 1, j), vel);", 0),
    ("advect_w_vertical_upwind_direction", "o_advect.c",
     "A3(w, i, k, j), A3(w, i, k + 1, j), -vel);", "A3(w, i, k, j), A3(w, i, k + 1, j), vel);", 0),
    ("advect_v_face_velocity_index", "o_advect.c",
     "real vel = (real)0.5 * (A3(rv, i, k, j) + A3(rv, i, k, j - 1));",
     "real vel = (real)0.5 * (A3(rv, i, k, j) + A3(rv, i, k, j + 1));", 0),
    ("advect_u_cross_velocity_index", "o_advect.c",
     "real vel = (real)0.5 * (A3(rv, i, k, j) + A3(rv, i - 1, k, j));",
     "real vel = (real)0.5 * (A3(rv, i, k, j) + A3(rv, i + 1, k, j));", 0),
    ("advect_w_x_divergence_sign", "o_advect.c",
     "for (int i = 2; i <= ide - 2; i++) A3(tend, i, k, j) -= A2(msftx, i, j) * rdx * (fqx[i + 1] - fqx[i]);",
     "for (int i = 2; i <= ide - 2; i++) A3(tend, i, k, j) += A2(msftx, i, j) * rdx * (fqx[i + 1] - fqx[i]);", 0),
    ("coriolis_v_sign", "o_bigstep.c", "val = val - (A2(msfvy, i, j) / A2(msfvx, i, j)) * (real)0.5 * (A2(f, i, j)",
     "val = val + (A2(msfvy, i, j) / A2(msfvx, i, j)) * (real)0.5 * (A2(f, i, j)", 0),
    ("curvature_w_sign", "o_bigstep.c", "rer * ((real)0.5 * (A1(fzm, k) * (A3(ru, i, k, j)",
     "-rer * ((real)0.5 * (A1(fzm, k) * (A3(ru, i, k, j)", 0),
    ("smag_prandtl_inverted", "o_diff.c", "A3(xkhh, i, k, j) = km / pr;", "A3(xkhh, i, k, j) = km * pr;", 0),
    ("smag_slope_branch_swapped", "o_diff.c", "real divisor = sd > lim ? alpha * alpha : alpha;",
     "real divisor = sd > lim ? alpha : alpha * alpha;", 0),
    ("hdiff_v_sign", "o_diff.c", "A3(tend, i, k, j) = A3(tend, i, k, j) + inc;", "A3(tend, i, k, j) = A3(tend, i, k, j) - inc;", 1),
    ("hdiff_w_sign", "o_diff.c", "A3(tend, i, k, j) = A3(tend, i, k, j) + inc;", "A3(tend, i, k, j) = A3(tend, i, k, j) - inc;", 2),
    ("hdiff_s_sign", "o_diff.c", "A3(tend, i, k, j) = A3(tend, i, k, j) + inc;", "A3(tend, i, k, j) = A3(tend, i, k, j) - inc;", 3),
    ("sumflux_no_substep_mean", "o_small.c", "if (fin) acc = acc / ns + ((A1(c1h, k) * A2(muu, i, j)",
     "if (fin) acc = acc + ((A1(c1h, k) * A2(muu, i, j)", 0),
    ("advance_uv_pressure_difference_order", "o_small.c",
     "* (A3(p, i, k, j) - A3(p, i - di, k, j - dj));", "* (A3(p, i - di, k, j - dj) - A3(p, i, k, j));", 0),
    ("advance_w_forward_sweep_sign", "o_small.c", "wl[k] = (wl[k] - A3(a, i, k, j) * wl[k - 1]) * A3(alpha, i, k, j);",
     "wl[k] = (wl[k] + A3(a, i, k, j) * wl[k - 1]) * A3(alpha, i, k, j);", 0),
    ("advance_w_backward_sweep_sign", "o_small.c", "wl[k] = wl[k] - A3(gamma, i, k, j) * wl[k + 1];",
     "wl[k] = wl[k] + A3(gamma, i, k, j) * wl[k + 1];", 0),
    ("coef_w_super_diagonal_level", "o_small.c", "A3(c2a, i, k, j)) / (mh * mfa);", "A3(c2a, i, k, j)) / (mh * mf);", 0),
    ("coef_w_top_pivot_sign", "o_small.c", "(real)1 / (btop - A3(a, i, kde, j) * gprev);", "(real)1 / (btop + A3(a, i, kde, j) * gprev);", 0),
    ("relax_weight_coefficient", "o_step.c", "(real)((real)0.1 / (real)cfg->dt)", "(real)((real)0.2 / (real)cfg->dt)", 0),
    ("n2_virtual_coefficient", "o_diff.c", "c161 = (real)1.61;", "c161 = (real)1.0;", 0),
    ("n2_saturated_theta_e_difference", "o_diff.c", "(thetaep1 - thetaem1)", "(thetaem1 - thetaep1)", 0),
    ("n2_dry_theta_one_sided", "o_diff.c", "(A3(th, i, k + 1, j) - A3(th, i, k - 1, j)) / thk",
     "(A3(th, i, k + 1, j) - A3(th, i, k, j)) / thk", 0),
    ("advect_scalar_x_divergence_sign", "o_advect.c",
     "for (int i = 2; i <= ide - 2; i++) A3(tend, i, k, j) -= A2(msftx, i, j) * rdx * (FQ(i + 1, k) - FQ(i, k));",
     "for (int i = 2; i <= ide - 2; i++) A3(tend, i, k, j) += A2(msftx, i, j) * rdx * (FQ(i + 1, k) - FQ(i, k));", 0),
]


def _plant(dst: str, fname: str, old: str, new: str, occurrence: int) -> None:
    path = os.path.join(dst, fname)
    text = open(path).read()
    pos = -1
    for _ in range(occurrence + 1):
        pos = text.find(old, pos + 1)
        if pos < 0:
            raise AssertionError(f"{fname}: occurrence {occurrence} of {old!r} not found")
    with open(path, "w") as fh:
        fh.write(text[:pos] + new + text[pos + len(old):])


def _build_and_run(root: str, mutation) -> tuple[str, int, str]:
    src = os.path.join(root, "src")
    out = os.path.join(root, "_build")
    shutil.copytree(SRC, src)
    os.makedirs(out)
    if mutation is not None:
        _plant(src, *mutation[1:])
    for prec, real in (("f32", "float"), ("f64", "double")):
        cmd = ["gcc", *CFLAGS, f"-DOREAL={real}", "-o", os.path.join(out, f"liboracle_{prec}.so"),
               *[os.path.join(src, f) for f in SOURCES], "-lm"]
        subprocess.run(cmd, check=True, capture_output=True)
    env = dict(os.environ, WRFDC_ORACLE_BUILD=out, OMP_NUM_THREADS="1")
    res = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", *PINS],
                         cwd=REPO, env=env, capture_output=True, text=True, timeout=600)
    name = "unmodified" if mutation is None else mutation[0]
    return name, res.returncode, res.stdout[-400:]


def test_every_planted_slip_fails_a_pin(tmp_path):
    cases = [None] + MUTATIONS
    workers = max(1, min(8, os.cpu_count() or 1))
    results = {}
    with cf.ThreadPoolExecutor(max_workers=workers) as pool:
        futs = [pool.submit(_build_and_run, str(tmp_path / f"m{n}"), m) for n, m in enumerate(cases)]
        for f in futs:
            name, rc, tail = f.result()
            results[name] = (rc, tail)
    rc, tail = results.pop("unmodified")
    assert rc == 0, f"the pins fail on the unmodified oracle:\n{tail}"
    survivors = {n: t for n, (rc, t) in results.items() if rc != 1}
    assert not survivors, "slips that no pin catches (or that broke the run):\n" + "\n".join(
        f"{n}: {t}" for n, t in survivors.items())


@pytest.mark.parametrize("mutation", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_mutation_text_present(mutation, tmp_path):
    """Each planted slip names text that occurs in the current sources."""
    shutil.copytree(SRC, tmp_path / "src")
    _plant(str(tmp_path / "src"), *mutation[1:])
