"""CPU: the planted-mistake check of the oracle, automated.

Each entry below is a plausible one-line mistake in the oracle's C source
(a wrong sign, a dropped term, a transposed index, a wrong face, a missing
weight).  For each, a copy of oracle/ is altered by exactly that edit, built
with the oracle's own Makefile in a temporary folder, and the oracle's pin
suites are run against the altered library (DSMC_ORACLE_LIB).  The check
passes only if every altered oracle fails at least one pin: a mistake the
pins cannot see would also pass the GPU parity tests if the device code
made it too, since those compare the kernels with this oracle."""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
PIN_SUITES = ["tests/test_oracle_pins.py", "tests/test_oracle_walls.py", "tests/test_oracle_step.py",
              "tests/test_oracle_rotation.py", "tests/test_oracle_amr.py"]

# (id, what the mistake is, exact text of the oracle, its replacement)
MUTATIONS = [
    ("inflow_face3_outward", "face y = L2 inflow given v_y = +v_n (outward)",
     "y = cs->L2;        vy = -vn;", "y = cs->L2;        vy = vn;"),
    ("inflow_face1_at_x0", "face x = L1 inflow placed at x = 0",
     "else if (f == 1) { x = cs->L1;", "else if (f == 1) { x = 0.0f;"),
    ("inflow_face2_no_drift", "face y = 0 inflow without its tangential drift V_jet",
     "vy = vn;  vx = cs->inj_ux + cs->inj_sig * g0; }", "vy = vn;  vx = cs->inj_sig * g0; }"),
    ("inflow_x_faces_no_vy_drift", "x-face inflow without the tangential drift V_y",
     "y = upos * cs->L2; vx = vn;  vy = cs->inj_uy + cs->inj_sig * g0; }",
     "y = upos * cs->L2; vx = vn;  vy = cs->inj_sig * g0; }"),
    ("inflow_full_step", "inflow particles fly the whole step instead of a random part",
     "move_one(cs, &x, &y, &vx, &vy, &vz, oer ? &er : NULL, pid, step, ufrac * cs->dt, dead, err);",
     "move_one(cs, &x, &y, &vx, &vy, &vz, oer ? &er : NULL, pid, step, cs->dt, dead, err);"),
    ("inflow_no_flux_weight", "inflow normal speed without the flux weight v_n",
     "float acc = (cand / ws) * or_expf(ref - d * d);", "float acc = or_expf(ref - d * d);"),
    ("specular_body_sign", "specular body reflection with the wrong sign",
     "vx = vx - 2.0f * vn * S[4];\n                vy = vy - 2.0f * vn * S[5];",
     "vx = vx + 2.0f * vn * S[4];\n                vy = vy + 2.0f * vn * S[5];"),
    ("first_segment_not_earliest", "the first segment hit in list order instead of the earliest",
     "                best = s;\n                kind = 4 + k;\n            }",
     "                best = s;\n                kind = 4 + k;\n                break;\n            }"),
    ("end_overlap_any_flight", "segment end overlap taken by every flight, not only one entering the neighbour",
     "if (!(vx * Q[4] + vy * Q[5] < 0.0f)) continue;", "if (0) continue;"),
    ("no_end_overlap", "segment ends without overlap (a ray through a vertex slips between the tests)",
     "if (u < -OR_SEG_TOL || u > 1.0f + OR_SEG_TOL) continue;", "if (u < 0.0f || u > 1.0f) continue;"),
    ("segment_reach_x10", "flights up to 10 cells test only their cell's segment list",
     "int all = reach > cs->margin;", "int all = reach > 10.0f * cs->margin;"),
    ("diffuse_vn_sigma", "diffuse normal speed scaled by sigma instead of v_mp",
     "float vn = sqrtf(-or_logf(u01o(w[0]))) * vmp;", "float vn = sqrtf(-or_logf(u01o(w[0]))) * sig;"),
    ("diffuse_vt_offset", "diffuse tangential component offset by 5 % of sigma",
     "float vt = r * c;", "float vt = r * c + 0.05f * sig;"),
    ("diffuse_frame_sign", "diffuse tangential direction along +n rotated the wrong way",
     "*vx = vn * nx + vt * (-ny);", "*vx = vn * nx + vt * ny;"),
    ("face_specular_axis", "specular faces flip the tangential component",
     "if (kind < 2) vx = -vx; else vy = -vy;", "if (kind < 2) vy = -vy; else vx = -vx;"),
    ("scatter_cos_U", "scattering cosine U instead of 2U - 1",
     "float cosr = 2.0f * u01(w[3]) - 1.0f;", "float cosr = u01(w[3]);"),
    ("scatter_azimuth", "both transverse components of g' from the cosine of phi",
     "rz = g * (sinr * sp);", "rz = g * (sinr * cp);"),
    ("collision_cm", "post-collision centre of mass from particle 1 alone",
     "float mx = 0.5f * (vx[p1] + vx[p2]);", "float mx = vx[p1];"),
    ("ntc_np_squared", "NTC candidates from np^2 instead of np(np - 1)",
     "(float)((long long)np * (long long)(np - 1)) * cs->ntc_coef;",
     "(float)((long long)np * (long long)np) * cs->ntc_coef;"),
    ("ntc_no_remainder", "NTC candidate remainder dropped",
     "float tot = e + *mrem;", "float tot = e;"),
    ("vhs_exponent_sign", "VHS cross-section exponent with the wrong sign",
     "cs->K1 * or_expf(cs->e_exp * or_logf(g2))", "cs->K1 * or_expf(-cs->e_exp * or_logf(g2))"),
    ("sample_no_vz2", "sampled v^2 without the vz term",
     "s2 += a * a + b * b + d * d;", "s2 += a * a + b * b;"),
    ("cell_index_transposed", "cell index i + j*nx instead of j + i*ny",
     "return j + i * cs->ny;", "return i + j * cs->nx;"),
    ("scan_inclusive", "inclusive instead of exclusive scan",
     "for (int c = 0; c < ncell; ++c) { start[c] = s; s += count[c]; }",
     "for (int c = 0; c < ncell; ++c) { s += count[c]; start[c] = s; }"),
    ("fill_vy_cosine", "fill v_y from the cosine of the v_x pair (v_x, v_y correlated)",
     "vy[i] = uy[c] + sig[c] * (r1 * s1);", "vy[i] = uy[c] + sig[c] * (r1 * c1);"),
    # rotational energy (zeta = 2)
    ("lb_full_mass", "collision energy with the full mass in E_t (the reference's bookkeeping)",
     "float ec = 0.25f * g2 + (er[p1] + er[p2]);", "float ec = 0.5f * g2 + (er[p1] + er[p2]);"),
    ("lb_beta_one_factor", "translational share Beta(5/2 - omega, 1): one factor of the product dropped",
     "float x = or_expf(cs->lb_ia * or_logf(u01o(w3[1])) + cs->lb_ib * or_logf(u01o(w3[2])));",
     "float x = or_expf(cs->lb_ia * or_logf(u01o(w3[1])));"),
    ("lb_split_half", "rotational energy halved between the two (the reference's live code)",
     "float e1 = erot * u01(w3[3]);", "float e1 = erot * 0.5f;"),
    ("lb_rot_prob_ignored", "every collision exchanges whatever rot_prob is",
     "if (u01(w3[0]) < cs->rot_p) {", "if (1) {"),
    ("lb_speed_kept", "relative speed not rescaled to the new translational energy",
     "g = 2.0f * sqrtf(et);", "(void)et;"),
    ("wall_rot_not_accommodated", "diffuse walls keep the rotational energy",
     "if (er) *er = rot_energy(w[3], ekt);", "(void)ekt;"),
    ("inflow_rot_wall_temperature", "inflow rotational energy at the wall temperature",
     "er = rot_energy(w[0], cs->inj_ekt);", "er = rot_energy(w[0], cs->body_ekt);"),
    ("fill_rot_per_dof", "fill rotational energy at kT/2 (one degree of freedom)",
     "er[i] = rot_energy(w[0], sig[c] * sig[c]);", "er[i] = rot_energy(w[0], 0.5f * sig[c] * sig[c]);"),
    ("amr_leaf_volume_of_cell", "sub-cell candidates with the cell's volume instead of the leaf's",
     "float vleaf = vol / (float)nl;", "float vleaf = vol;"),
    ("amr_pairs_across_cell", "sub-cell pairs drawn from the whole cell",
     "nc += ntc_candidates(cs, c, step, s0 + ls[l], npl, m0, M, perm, vx, vy, vz, er, &sm);",
     "nc += ntc_candidates(cs, c, step, s0, np, m0, M, perm, vx, vy, vz, er, &sm);"),
    ("amr_no_leaf_remainder", "sub-cell candidate remainder not carried from leaf to leaf",
     "int M = ntc_count(cs, npl, vleaf, smax0, &mrem[c]);",
     "float r0 = 0.0f; int M = ntc_count(cs, npl, vleaf, smax0, &r0);"),
    ("amr_grid_floor", "sub-cell count rounded down instead of up", "double n = ceil(q);", "double n = floor(q);"),
    ("amr_level3_from_largest", "level 3 sized by the largest mean free path",
     "double mn = had ? lam_min : 2.0 * l2x;", "double mn = had ? lam_max : 2.0 * l2x;"),
    ("amr_mfp_exponent_sign", "mean free path temperature exponent with the wrong sign",
     "float f = or_expf((float)(g->omega - 0.5) * or_logf(tr));",
     "float f = or_expf((float)(0.5 - g->omega) * or_logf(tr));"),
    ("amr_leaf_transposed", "leaf index with the axes transposed",
     "    return ly + lx * nsy;\n}", "    return lx + ly * nsx;\n}"),
    ("amr_vts_no_flow_speed", "VTS time step without the flow speed",
     "dt[c] = 0.5 * l / (sqrt(uu) + sqrt(2.0 * g->boltz * T / g->mass));",
     "dt[c] = 0.5 * l / sqrt(2.0 * g->boltz * T / g->mass);"),
]


def _build(tmp, mid, old, new):
    d = os.path.join(tmp, mid)
    os.makedirs(d)
    for f in ("dsmc_oracle.c", "dsmc_oracle.h", "Makefile"):
        shutil.copy(os.path.join(ORACLE, f), d)
    src = os.path.join(d, "dsmc_oracle.c")
    text = open(src).read()
    if old:
        n = text.count(old)
        if n != 1:
            return None, f"pattern found {n} times"
        open(src, "w").write(text.replace(old, new))
    r = subprocess.run(["make", "-s", "-C", d], capture_output=True, text=True)
    if r.returncode != 0:
        return None, "build failed: " + r.stderr[-400:]
    return os.path.join(d, "liboracle.so"), ""


def _run_pins(lib):
    env = dict(os.environ, DSMC_ORACLE_LIB=lib, OMP_THREAD_LIMIT="1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", "-m", "not gpu",
                        *PIN_SUITES], cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    return r.returncode, r.stdout[-600:]


def _check(tmp, m):
    mid, what, old, new = m
    lib, msg = _build(tmp, mid, old, new)
    if lib is None:
        return mid, "error", msg
    rc, out = _run_pins(lib)
    # pytest: 1 = a test failed (the mistake is caught), 0 = all passed
    return mid, {0: "missed", 1: "caught"}.get(rc, "error"), out


def test_mutation_table_is_well_formed():
    ids = [m[0] for m in MUTATIONS]
    assert len(ids) == len(set(ids))
    src = open(os.path.join(ORACLE, "dsmc_oracle.c")).read()
    for mid, _, old, new in MUTATIONS:
        assert src.count(old) == 1, mid
        assert old != new, mid


@pytest.mark.slow
def test_every_planted_mistake_fails_a_pin(tmp_path):
    # the same build and pin run without any edit must pass, or every
    # altered oracle would count as caught
    lib, msg = _build(str(tmp_path), "unaltered", "", "")
    assert lib is not None, msg
    rc, out = _run_pins(lib)
    assert rc == 0, "pins fail on the unaltered oracle:\n" + out
    with ThreadPoolExecutor(max_workers=max(2, min(8, os.cpu_count() or 2))) as ex:
        res = list(ex.map(lambda m: _check(str(tmp_path), m), MUTATIONS))
    bad = [(mid, st, out) for mid, st, out in res if st != "caught"]
    assert not bad, "\n".join(f"{mid}: {st}\n{out}" for mid, st, out in bad)
