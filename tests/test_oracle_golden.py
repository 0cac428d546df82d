"""Pins the CPU oracle against data the reference itself produced
(tests/golden, copied from Hydrograd.jl's examples by make_fixtures.py).

* S0_cells_truth / zb_cell_truth / ManningN_cells_truth / hstill_truth are
  written by postprocess_forward_simulation_results_swe_2D
  (process_forward_simulation_results_2D.jl:4-81) with the reference's own
  update_bed_data and set-up: they pin mesh geometry, node->cell bed, the
  Green-Gauss bed slope and the zone map BIT-EXACTLY.
* friction_x/y_truth come from the reference's compute_friction_terms on
  the truth state: they pin the Manning term (q is reconstructed from
  u_truth = q / (h + h_small), hence the 1e-15 tolerance).
* The truth states are Tsit5 solutions at t = 100 s (steady upstream of the
  bump's hydraulic jump; a slow backwater transient in the uniform-flow
  channel): the oracle's Euler run must reach them to the fixture's own
  tolerance (reltol 1e-3).
* sensitivity_results.json is ForwardDiff's dQ(t_end)/dn_zone: the oracle's
  Dual run pins the forward-mode derivative at steady state.
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, load_golden_case

import pyoracle

CASES = {
    "bump": dict(ic=(0.33, 0.2, 0.0, 0.0)),
    "uniform": dict(ic=(3.857205, 3.0, 0.0, 0.0)),
}


def savannah_ic_path(tmp_path):
    d = np.load(os.path.join(GOLDEN, "savannah", "ic.npz"))
    p = tmp_path / "ic.json"
    p.write_text(json.dumps({k: d[k].tolist() for k in d.files}))
    return str(p)


def load(name, tmp_path):
    if name == "savannah":
        return load_golden_case("savannah", ic=savannah_ic_path(tmp_path))
    return load_golden_case(name, **CASES[name])


@pytest.mark.parametrize("name", ["bump", "uniform", "savannah"])
def test_bed_slope_and_setup_bit_exact(name, tmp_path):
    c = load(name, tmp_path)
    t = np.load(os.path.join(GOLDEN, name, "truth.npz"))
    N = c.N
    assert np.array_equal(c.zb, t["zb_cell_truth"])
    assert np.array_equal(c.n, t["ManningN_cells_truth"])
    assert np.array_equal(c.hstill, t["hstill_truth"])
    S0 = t["S0_cells_truth"]                      # N x 2, column-major
    o = pyoracle.Oracle(c)
    _, sx, sy = o.bed_data(c.zb)
    assert np.array_equal(sx, S0[:N]) and np.array_equal(sy, S0[N:])
    assert np.abs(S0).max() > 0 or name == "uniform"


@pytest.mark.parametrize("name", ["bump", "uniform", "savannah"])
def test_friction_matches_reference_terms(name, tmp_path):
    c = load(name, tmp_path)
    t = np.load(os.path.join(GOLDEN, name, "truth.npz"))
    h = t["h_truth"]
    qx = t["u_truth"] * (h + c.h_small)
    qy = t["v_truth"] * (h + c.h_small)
    fx, fy = pyoracle.friction(h, qx, qy, c.n)
    for f, ft in ((fx, t["friction_x_truth"]), (fy, t["friction_y_truth"])):
        scale = np.abs(t["friction_x_truth"]).max()
        assert np.max(np.abs(f - ft)) <= 1e-14 * scale


def test_geometry_invariants(tmp_path):
    c = load("savannah", tmp_path)
    m = c.mesh
    # closed polygons: sum_k n_k L_k = 0 per cell
    off = m.face_off.astype(np.int64)
    sx = np.add.reduceat(m.face_nx * m.face_len, off[:-1])
    sy = np.add.reduceat(m.face_ny * m.face_len, off[:-1])
    per = np.add.reduceat(m.face_len, off[:-1])
    assert np.max(np.abs(sx) / per) < 1e-13 and np.max(np.abs(sy) / per) < 1e-13
    # unit normals, outward normal of a boundary face == the cell's own normal
    assert np.max(np.abs(m.face_nx ** 2 + m.face_ny ** 2 - 1.0)) < 1e-15
    assert np.array_equal(m.face_nx[m.ghost_slot], m.ghost_nx)
    assert np.array_equal(m.face_ny[m.ghost_slot], m.ghost_ny)
    # pairing: n_ij = -n_ji exactly
    cell_of_slot = np.repeat(np.arange(c.N), m.cell_deg)
    for s in range(0, m.face_off[-1], 7):
        j = m.face_nbr[s]
        if j < 0:
            continue
        i = cell_of_slot[s]
        t = [u for u in range(m.face_off[j], m.face_off[j + 1]) if m.face_nbr[u] == i][0]
        assert m.face_nx[t] == -m.face_nx[s] and m.face_ny[t] == -m.face_ny[s]
        assert m.face_len[t] == m.face_len[s]


# (case, T, dt, upstream max tol, h L1 tol, u L1 tol, steady residual checked)
EULER_CASES = [("bump", 100.0, 0.01, 1e-5, 2e-3, 5e-3, True),
               ("uniform", 100.0, 0.004, 1e-3, 1e-3, 1e-2, False)]


@pytest.mark.parametrize("name,T,dt,up_tol,h_tol,u_tol,steady", EULER_CASES)
def test_euler_reaches_reference_state(name, T, dt, up_tol, h_tol, u_tol, steady, tmp_path):
    c = load(name, tmp_path)
    t = np.load(os.path.join(GOLDEN, name, "truth.npz"))
    o = pyoracle.Oracle(c)
    Q = o.run(c.Q0, dt, int(round(T / dt)))
    h = Q[: c.N] + c.hstill
    x = c.mesh.centroid[:, 0]
    # upstream of the bump crest the flow is subcritical and steady
    up = x < 10.0
    assert np.max(np.abs(h[up] - t["h_truth"][up])) / np.max(t["h_truth"]) < up_tol
    # whole domain (the bump's hydraulic jump is resolved with odd/even
    # wiggles by the Tsit5 run; the uniform channel is still a transient whose
    # Euler error is O(dt)), L1 against the fixture's reltol 1e-3
    assert np.mean(np.abs(h - t["h_truth"])) / np.mean(t["h_truth"]) < h_tol
    u = Q[c.N: 2 * c.N] / (h + c.h_small)
    assert np.mean(np.abs(u - t["u_truth"])) / np.mean(np.abs(t["u_truth"])) < u_tol
    if steady:  # RHS of the reference's own steady state vanishes against the flux scale
        Qt = np.concatenate([t["xi_truth"], t["u_truth"] * (t["h_truth"] + c.h_small),
                             t["v_truth"] * (t["h_truth"] + c.h_small)])
        r = o.rhs(Qt)
        scale = 0.5 * c.g * np.max(t["h_truth"]) ** 2 / np.sqrt(np.min(c.mesh.area))
        assert np.max(np.abs(r.reshape(3, c.N)[:, up])) / scale < 1e-5
        assert np.mean(np.abs(r)) / scale < 1e-3


def _sens_case(name):
    with open(os.path.join(GOLDEN, name, "run_control.json")) as fh:
        txt = fh.read()
    rc = json.loads(txt)
    sa = rc["sensitivity_analysis_options"]
    ic = tuple(sa["sensitivity_forward_simulation_initial_condition_constant_values"])
    zn = np.asarray(sa["sensitivity_ManningN_values"], dtype=np.float64)
    c = load_golden_case(name, ic=ic)
    T = rc["time_settings"]["tspan"][1]
    return c, zn, T


def _sens_errors(name, dt):
    c, zn, T = _sens_case(name)
    S = np.load(os.path.join(GOLDEN, name, "sens.npz"))["sensitivity_results"].reshape(len(zn), 3 * c.N).T
    c = c.copy_fields(n=zn[c.zone])
    o = pyoracle.Oracle(c)
    N = c.N
    errs = []
    for z in range(len(zn)):
        dn = (c.zone == z).astype(np.float64)
        _, dQ, _, _ = o.jvp_run(c.Q0, dt, int(round(T / dt)), dn=dn)
        ref = S[:, z]
        if not np.any(dn):  # empty zone: ForwardDiff column is exactly zero
            assert np.max(np.abs(ref)) == 0.0 and np.max(np.abs(dQ)) == 0.0
            continue
        for comp in range(2):  # xi and q_x (q_y ~ 0 in these one-cell-wide channels)
            a, b = dQ[comp * N:(comp + 1) * N], ref[comp * N:(comp + 1) * N]
            errs.append(np.max(np.abs(a - b)) / np.max(np.abs(b)))
    return max(errs)


def test_forward_mode_matches_forwarddiff_at_steady_state():
    """dQ(t_end)/dn_zone of the oracle's Dual run vs ForwardDiff through Tsit5
    (sensitivity_results.json, 3N x nzones column-major) on the bump case,
    which is steady at t_end = 200 s: integrator-independent, so it must
    agree to the fixture's own tolerance."""
    assert _sens_errors("bump_sens", 0.02) < 1e-3


def test_forward_mode_converges_to_forwarddiff_transient():
    """The uniform-flow channel is still a transient at t_end = 100 s: the
    Euler derivative differs from Tsit5's by O(dt), so halving dt must halve
    the gap (first-order convergence toward the fixture)."""
    e1 = _sens_errors("uniform_sens", 0.004)
    e2 = _sens_errors("uniform_sens", 0.002)
    assert e1 < 0.1
    assert 0.4 < e2 / e1 < 0.6
