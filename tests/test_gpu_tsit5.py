"""GPU checks of the adaptive Tsit5 integrator over the device RHS.

* Against the reference's own 101-frame Tsit5 trajectories (forward pass of
  the ManningN sensitivity examples, OrdinaryDiffEq Tsit5 at its default
  tolerances): every saved frame, transient included.  Integrator- and
  tolerance-dependent, so the bar is the fixtures' own accuracy.
* Against the explicit Euler path of the same RHS at small dt (first-order
  convergence toward the Tsit5 solution).
"""
import json
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, load_golden_case
from hgswe import capi, cases

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def sens_case(name):
    with open(os.path.join(GOLDEN, name, "run_control.json")) as fh:
        rc = json.load(fh)
    sa = rc["sensitivity_analysis_options"]
    ic = tuple(sa["sensitivity_forward_simulation_initial_condition_constant_values"])
    zn = np.asarray(sa["sensitivity_ManningN_values"], dtype=np.float64)
    c = load_golden_case(name, ic=ic)
    c = c.copy_fields(n=zn[c.zone])
    T = rc["time_settings"]["tspan"][1]
    traj = np.load(os.path.join(GOLDEN, name, "traj.npz"))["forward_simulation_results"]
    frames = traj.reshape(-1, 3 * c.N)  # column-major 3N x nframes in the reference
    return c, T, frames


# The fixtures were integrated at OrdinaryDiffEq's defaults (reltol 1e-3), so
# single frames of the fast transient carry up to a few 1e-3 of their own
# integration error; the median frame is far tighter.
@pytest.mark.parametrize("name,tol,tol_median", [("bump_sens", 2e-3, 1e-4), ("uniform_sens", 1e-2, 1e-3)])
def test_tsit5_matches_reference_trajectory(name, tol, tol_median):
    c, T, frames = sens_case(name)
    nf = frames.shape[0]
    ctx = capi.Context(c)
    Q, st = ctx.tsit5(c.Q0, 0.0, T, nf - 1, abstol=1e-9, reltol=1e-7, dt0=1e-3)
    N = c.N
    errs = []
    for f in range(nf):
        e = []
        for k in range(2):  # xi and q_x (q_y ~ 0 in these one-cell-wide channels)
            a, b = Q[f, k * N:(k + 1) * N], frames[f, k * N:(k + 1) * N]
            e.append(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-12))
        errs.append(max(e))
    errs = np.array(errs)
    say(f"  {name}: steps {st}, frame rel err max {errs.max():.2e} (frame {int(errs.argmax())}), "
        f"median {np.median(errs):.2e}, final {errs[-1]:.2e}")
    assert errs[0] < 1e-15
    assert errs.max() < tol
    assert np.median(errs) < tol_median


def test_euler_converges_to_tsit5():
    c = cases.dam_break(nx=60, ny=4, nsteps=0, h_left=1.0, h_right=0.6)
    T = 2.0
    ctx = capi.Context(c)
    Q, st = ctx.tsit5(c.Q0, 0.0, T, 1, abstol=1e-10, reltol=1e-9, dt0=1e-3)
    ref = Q[-1]
    e = []
    for dt in (0.004, 0.002):
        q = ctx.step(c.Q0, dt, int(round(T / dt)))
        e.append(np.max(np.abs(q - ref)) / np.max(np.abs(ref)))
    say(f"  Euler vs Tsit5 at t={T}: dt 0.004 -> {e[0]:.3e}, dt 0.002 -> {e[1]:.3e} (steps {st})")
    assert e[1] < e[0] and 0.35 < e[1] / e[0] < 0.65
