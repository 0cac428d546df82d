"""TEST INFRASTRUCTURE ONLY: NumPy restatement of Hydrograd.jl's variable
Manning-n laws (the forward simulation's "variable" ManningN option), used
as the checker for the library's device laws.  Only the test-suite may
import this module; the product path never does.

Each law takes the water depth h, the velocity magnitude Umag and the
roughness height ks per cell and returns (n, h_ks, f, Re) like the
reference's update_ManningN_forward_simulation
(src/parameters/process_ManningN_2D.jl:102-117); the three depth-only laws
return zero h_ks, f and Re (:143-145, :158-160, :175-177).  Evaluation order
follows the reference expression by expression (NumPy evaluates left to
right, without fused multiply-add).

Pins (tests/test_nlaw_oracle.py): the reference's own final-state fixtures
forward_simulation/oneD_channel_with_bump_ManningN_h (sigmoid law) and
forward_simulation/Savannah_River_ManningN_ks_h_Umag (h_Umag_ks law) hold
ManningN_cells_truth, h_ks_cells_truth, friction_factor_cells_truth and
Re_cells_truth evaluated at their h_truth / u_truth / v_truth
(process_forward_simulation_results_2D.jl:27-38); closed forms (each
depth-only law at its limits; the friction factor at the laminar limit
f = 24 / Re) cover the power and inverse laws that no fixture holds.

The state features the RHS feeds the law (semi_discretize_swe_2D.jl:101-106,
130-134) are in rhs_features.
"""
from __future__ import annotations

import numpy as np

EPS = np.finfo(np.float64).eps

POWER, SIGMOID, INVERSE, H_UMAG_KS = 1, 2, 3, 4


def inverse_n(h, n_lower, n_upper, k):
    """n = n_lower + (n_upper - n_lower) / (1 + k h)  (process_ManningN_2D.jl:139-150)."""
    z = np.zeros_like(h)
    return n_lower + (n_upper - n_lower) / (1.0 + k * h), z, z.copy(), z.copy()


def power_law_n(h, n_lower, n_upper, k):
    """n = n_lower + (n_upper - n_lower) (h + eps)^-k  (process_ManningN_2D.jl:154-165)."""
    z = np.zeros_like(h)
    return n_lower + (n_upper - n_lower) * np.power(h + EPS, -k), z, z.copy(), z.copy()


def sigmoid_n(h, n_lower, n_upper, k, h_mid):
    """n = n_lower + (n_upper - n_lower) / (1 + exp(k (h - h_mid)))  (process_ManningN_2D.jl:170-182)."""
    z = np.zeros_like(h)
    return n_lower + (n_upper - n_lower) / (1.0 + np.exp(k * (h - h_mid))), z, z.copy(), z.copy()


def manning_n_h_Umag_ks(h, Umag, ks):
    """Cheng (2008) friction factor and its Manning n (process_ManningN_2D.jl:186-213):
    Re = Umag h / nu, alpha = 1 / (1 + (Re/850)^9), beta = 1 / (1 + (Re / (160 h/ks))^2),
    f = 1 / ((Re/24)^alpha (1.8 log10(Re/2.1))^(2(1-alpha)beta) (2 log10(11.8 h/ks))^(2(1-alpha)(1-beta))),
    n = sqrt(f/8) h^(1/6) / sqrt(9.81)."""
    nu = 1.0e-6
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        Re = Umag * h / nu
        h_ks = h / ks
        alpha = 1.0 / (1.0 + np.power(Re / 850.0, 9))
        r = Re / (h_ks * 160.0)
        beta = 1.0 / (1.0 + r * r)
        part1 = np.power(Re / 24.0, alpha)
        part2 = np.power(1.8 * np.log10(Re / 2.1), 2.0 * (1.0 - alpha) * beta)
        part3 = np.power(2.0 * np.log10(11.8 * h_ks), 2.0 * (1.0 - alpha) * (1.0 - beta))
        f = 1.0 / (part1 * part2 * part3)
        n = np.sqrt(f / 8.0) * np.power(h, 1.0 / 6.0) / np.sqrt(9.81)
    return n, h_ks, f, Re


def evaluate(kind, h, Umag, ks=None, n_lower=0.0, n_upper=0.0, k=1.0, h_mid=1.0):
    """create_manning_function's dispatch (process_ManningN_2D.jl:122-134)."""
    h = np.asarray(h, dtype=np.float64)
    if kind == POWER:
        return power_law_n(h, n_lower, n_upper, k)
    if kind == SIGMOID:
        return sigmoid_n(h, n_lower, n_upper, k, h_mid)
    if kind == INVERSE:
        return inverse_n(h, n_lower, n_upper, k)
    if kind == H_UMAG_KS:
        return manning_n_h_Umag_ks(h, np.asarray(Umag, dtype=np.float64), np.asarray(ks, dtype=np.float64))
    raise ValueError(f"unknown Manning n law {kind}")


def rhs_features(Q, hstill, h_small):
    """The law's inputs inside swe_2d_rhs (semi_discretize_swe_2D.jl:101-106, 130-134):
    h = xi + hstill clamped to h_small (the discharge zeroed there), u = q_x / h,
    v = q_y / h, Umag = sqrt(u^2 + v^2)."""
    N = hstill.size
    xi, qx, qy = Q[:N], Q[N:2 * N], Q[2 * N:]
    h = xi + hstill
    dry = h <= h_small
    h = np.where(dry, h_small, h)
    qx = np.where(dry, 0.0, qx)
    qy = np.where(dry, 0.0, qy)
    u = qx / h
    v = qy / h
    return h, np.sqrt(u * u + v * v)


def rhs_n(kind, Q, hstill, h_small, ks=None, **params):
    """Manning n per cell that one RHS evaluation uses under the law."""
    h, U = rhs_features(Q, hstill, h_small)
    return evaluate(kind, h, U, ks, **params)[0]
