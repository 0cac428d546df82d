"""Pins of the oracle's variable Manning-n laws (oracle/nlaw.py).

* The reference's final-state fixtures: the sigmoid n(h) bump channel and the
  Savannah River h_Umag_ks case hold n (and h/ks, f, Re) evaluated at their
  final h, u, v (process_forward_simulation_results_2D.jl:27-38).
* Closed forms of each law: the depth-only laws at their limits and
  midpoints; the friction factor's laminar (f = 24/Re), smooth-turbulent
  ((1.8 log10(Re/2.1))^-2) and fully rough ((2 log10(11.8 h/ks))^-2) limits
  of Cheng (2008), and n = sqrt(f/8) h^(1/6) / sqrt(g).
* The RHS features: the dry clamp and u = q/h.
"""
import json
import os

import numpy as np

import nlaw
from conftest import GOLDEN, load_golden_case


def _truth(name):
    return np.load(os.path.join(GOLDEN, name, "truth.npz"))


def _options(name):
    with open(os.path.join(GOLDEN, name, "run_control.json")) as fh:
        return json.load(fh)["forward_simulation_options"]


def _umag(t):
    u, v = t["u_truth"], t["v_truth"]
    return np.sqrt(u * u + v * v)


def test_sigmoid_law_matches_bump_fixture():
    t = _truth("bump_nh")
    o = _options("bump_nh")
    assert o["forward_simulation_ManningN_function_type"] == "sigmoid"
    p = {k: float(v) for k, v in o["forward_simulation_ManningN_function_parameters"].items()}
    n, hks, f, Re = nlaw.evaluate(nlaw.SIGMOID, t["h_truth"], _umag(t), **p)
    assert np.array_equal(n, t["ManningN_cells_truth"])
    assert not hks.any() and not f.any() and not Re.any()


def savannah_ks_cells(c):
    """ks per cell: the run_control list indexed by material id
    (process_SRH_2D_input.jl:159-166, process_ManningN_2D.jl:56-61)."""
    ks = np.asarray(_options("savannah_ks")["forward_simulation_ManningN_function_parameters"]["ks"], dtype=np.float64)
    return ks[c.zone]


def test_h_umag_ks_law_matches_savannah_fixture(tmp_path):
    from test_oracle_golden import savannah_ic_path

    c = load_golden_case("savannah", ic=savannah_ic_path(tmp_path))
    t = _truth("savannah_ks")
    ks = savannah_ks_cells(c)
    n, hks, f, Re = nlaw.evaluate(nlaw.H_UMAG_KS, t["h_truth"], _umag(t), ks)
    for got, key in ((n, "ManningN_cells_truth"), (hks, "h_ks_cells_truth"), (f, "friction_factor_cells_truth"),
                     (Re, "Re_cells_truth")):
        ref = t[key]
        rel = float(np.max(np.abs(got - ref) / np.abs(ref)))
        assert rel <= 2e-15, (key, rel)


def test_depth_only_laws_closed_forms():
    lo, hi, k = 0.02, 0.05, 3.0
    # power law: (h + eps)^-k = 1 at h + eps = 1; n -> n_lower for deep water
    n = nlaw.evaluate(nlaw.POWER, np.array([1.0 - nlaw.EPS, 1e8]), None, n_lower=lo, n_upper=hi, k=k)[0]
    assert n[0] == hi and abs(n[1] - lo) < 1e-20
    # inverse law: n_upper at h = 0, the midpoint at k h = 1, n_lower far away
    n = nlaw.evaluate(nlaw.INVERSE, np.array([0.0, 1.0 / k, 1e12]), None, n_lower=lo, n_upper=hi, k=k)[0]
    assert n[0] == hi and abs(n[1] - 0.5 * (lo + hi)) < 1e-17 and abs(n[2] - lo) < 1e-12
    # sigmoid: midpoint at h_mid, n_upper for h << h_mid, n_lower for h >> h_mid
    n = nlaw.evaluate(nlaw.SIGMOID, np.array([0.3, -5.0, 5.0]), None, n_lower=lo, n_upper=hi, k=100.0,
                      h_mid=0.3)[0]
    assert n[0] == lo + (hi - lo) / 2.0 and abs(n[1] - hi) < 1e-17 and abs(n[2] - lo) < 1e-17


def test_friction_factor_limits():
    g = 9.81
    # laminar: Re = 100 (alpha = 1 - 4e-9): f = 24 / Re
    h = np.array([1e-4]); U = np.array([1.0]); ks = np.array([1e-3])
    n, hks, f, Re = nlaw.manning_n_h_Umag_ks(h, U, ks)
    assert abs(Re[0] - 100.0) < 1e-12 and abs(hks[0] - 0.1) < 1e-16
    assert abs(f[0] * Re[0] / 24.0 - 1.0) < 1e-6
    # smooth turbulent: alpha -> 0 (Re = 1e5), beta -> 1 (Re << 160 h/ks)
    h = np.array([0.1]); U = np.array([1.0]); ks = np.array([1e-7])
    n, hks, f, Re = nlaw.manning_n_h_Umag_ks(h, U, ks)
    assert abs(f[0] * (1.8 * np.log10(Re[0] / 2.1)) ** 2 - 1.0) < 1e-5
    # fully rough: alpha -> 0, beta -> 0 (Re >> 160 h/ks)
    h = np.array([10.0]); U = np.array([100.0]); ks = np.array([1.0])
    n, hks, f, Re = nlaw.manning_n_h_Umag_ks(h, U, ks)
    fr = (2.0 * np.log10(11.8 * 10.0)) ** -2
    assert abs(f[0] / fr - 1.0) < 1e-9
    # Manning n of f
    assert abs(n[0] - np.sqrt(f[0] / 8.0) * 10.0 ** (1.0 / 6.0) / np.sqrt(g)) < 1e-17


def test_rhs_features_dry_clamp():
    hstill = np.array([0.0, 0.0, 1.0])
    Q = np.array([2.0, 5e-4, -0.5, 4.0, 1.0, 7.0, 3.0, 1.0, 9.0])  # xi, q_x, q_y of 3 cells
    h, U = nlaw.rhs_features(Q, hstill, 1e-3)
    assert np.array_equal(h, [2.0, 1e-3, 0.5])
    assert U[0] == np.sqrt(2.0 ** 2 + 1.5 ** 2) and U[1] == 0.0 and U[2] == np.sqrt(14.0 ** 2 + 18.0 ** 2)
