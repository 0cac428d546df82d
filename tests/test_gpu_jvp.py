"""GPU checks of the forward-mode (JVP) path hg_jvp, the device counterpart
of ForwardDiff.Dual through the Euler step.

* Against the oracle's Dual run along the same tangent (per-cell n, bed zb
  with the bed-data rebuild, inlet discharge): 1e-9 relative after 60 steps.
* Against ForwardDiff's own dQ(t_end)/dn_zone fixtures (the reference's
  ManningN sensitivity example on the bump case, steady at t_end): to the
  fixture's tolerance, like the oracle's Dual run.
* The initial-state tangent against the oracle's Dual run.
"""
import json
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, load_golden_case
from hgswe import capi, cases

import pyoracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def channel():
    return cases.channel(6_000, seed=21, Lx=1500.0, Ly=150.0, Q_in=60.0, nsteps=60)


@pytest.mark.parametrize("kind", ["n", "zb", "Q"])
def test_jvp_matches_oracle_dual(channel, kind):
    c = channel
    nsteps = 60
    rng = np.random.default_rng(5)
    ctx = capi.Context(c)
    orc = pyoracle.Oracle(c)
    if kind == "n":
        ctx.set_active_param(capi.HG_PARAM_N_CELL)
        d = rng.normal(size=c.N) * 1e-3
        QN, dQN = ctx.jvp(c.Q0, c.dt, nsteps, params=c.n, dparams=d)
        Qo, dQo, _, _ = orc.jvp_run(c.Q0, c.dt, nsteps, dn=d)
    elif kind == "zb":
        ctx.set_active_param(capi.HG_PARAM_ZB)
        d = rng.normal(size=c.N) * 1e-2
        QN, dQN = ctx.jvp(c.Q0, c.dt, nsteps, params=c.zb, dparams=d)
        Qo, dQo, _, _ = orc.jvp_run(c.Q0, c.dt, nsteps, dzb=d)
    else:
        ctx.set_active_param(capi.HG_PARAM_Q)
        d = np.ones(c.mesh.n_inlet)
        QN, dQN = ctx.jvp(c.Q0, c.dt, nsteps, params=c.inletQ, dparams=d)
        Qo, dQo, _, _ = orc.jvp_run(c.Q0, c.dt, nsteps, dQin=d)
    say(f"  jvp [{kind}]: state rel {rel(QN, Qo):.2e}, tangent rel {rel(dQN, dQo):.2e}")
    assert rel(QN, Qo) <= 1e-10
    assert rel(dQN, dQo) <= 1e-9


def test_jvp_initial_state_tangent(channel):
    """The initial-state tangent against the oracle's Dual run (an exact
    derivative: difference quotients straddle the kink of the smoothed
    |lambda| = sqrt(u^2 + eps) at the fluid at rest)."""
    c = channel
    nsteps = 40
    rng = np.random.default_rng(6)
    v = np.concatenate([rng.normal(size=c.N) * 1e-3, rng.normal(size=2 * c.N) * 1e-3])
    QN, dQN = capi.Context(c).jvp(c.Q0, c.dt, nsteps, dQ0=v)
    Qo, dQo = pyoracle.Oracle(c).jvp_state(c.Q0, v, c.dt, nsteps)
    say(f"  jvp [Q0]: state rel {rel(QN, Qo):.2e}, tangent rel {rel(dQN, dQo):.2e}")
    assert rel(QN, Qo) <= 1e-10
    assert rel(dQN, dQo) <= 1e-9


def test_jvp_matches_forwarddiff_fixture():
    """dQ(t_end)/dn_zone of the device Dual run vs ForwardDiff through Tsit5
    (sensitivity_results.json) on the bump case, steady at t_end = 200 s."""
    name = "bump_sens"
    with open(os.path.join(GOLDEN, name, "run_control.json")) as fh:
        rc = json.load(fh)
    sa = rc["sensitivity_analysis_options"]
    ic = tuple(sa["sensitivity_forward_simulation_initial_condition_constant_values"])
    zn = np.asarray(sa["sensitivity_ManningN_values"], dtype=np.float64)
    c = load_golden_case(name, ic=ic)
    T = rc["time_settings"]["tspan"][1]
    S = np.load(os.path.join(GOLDEN, name, "sens.npz"))["sensitivity_results"].reshape(len(zn), 3 * c.N).T
    ctx = capi.Context(c)
    ctx.set_active_param(capi.HG_PARAM_N_ZONE, zone=c.zone, n_zones=len(zn))
    dt = 0.02
    N = c.N
    errs = []
    for z in range(len(zn)):
        dz = np.zeros(len(zn))
        dz[z] = 1.0
        _, dQ = ctx.jvp(c.Q0, dt, int(round(T / dt)), params=zn, dparams=dz)
        ref = S[:, z]
        if not np.any(c.zone == z):
            assert np.max(np.abs(ref)) == 0.0 and np.max(np.abs(dQ)) == 0.0
            continue
        for comp in range(2):
            a, b = dQ[comp * N:(comp + 1) * N], ref[comp * N:(comp + 1) * N]
            errs.append(np.max(np.abs(a - b)) / np.max(np.abs(b)))
    say(f"  jvp vs ForwardDiff fixture: max rel {max(errs):.2e}")
    assert max(errs) < 1e-3


def test_savannah_sensitivities_match_forwarddiff_fixture(tmp_path):
    """dQ(200 s)/dn_zone on the Savannah River mesh (1306 mixed cells, 6
    zones): the device Dual run and the device adjoint against ForwardDiff
    through Tsit5 (sensitivity_analysis/ManningN/Savana_River).  The Euler
    runs converge to the fixture at first order in dt (L1 error halves with
    dt); the max-norm error sits at the fixture's own reltol 1e-3 level."""
    from test_oracle_golden import savannah_ic_path

    with open(os.path.join(GOLDEN, "savannah_sens", "run_control.json")) as fh:
        rc = json.load(fh)
    zn = np.asarray(rc["sensitivity_analysis_options"]["sensitivity_ManningN_values"], dtype=np.float64)
    c = load_golden_case("savannah", ic=savannah_ic_path(tmp_path))
    T = rc["time_settings"]["tspan"][1]
    S = np.load(os.path.join(GOLDEN, "savannah_sens", "sens.npz"))["sensitivity_results"].reshape(len(zn), 3 * c.N).T
    N = c.N
    ctx = capi.Context(c)
    ctx.set_active_param(capi.HG_PARAM_N_ZONE, zone=c.zone, n_zones=len(zn))
    l1 = {}
    for dt in (0.05, 0.02):
        nsteps = int(round(T / dt))
        e1, emax = [], []
        J = np.zeros((3 * N, len(zn)))
        for z in range(len(zn)):
            dz = np.zeros(len(zn))
            dz[z] = 1.0
            _, J[:, z] = ctx.jvp(c.Q0, dt, nsteps, params=zn, dparams=dz)
            ref = S[:, z]
            if not np.any(c.zone == z):
                assert np.max(np.abs(ref)) == 0.0 and np.max(np.abs(J[:, z])) == 0.0
                continue
            for k in range(3):
                a, b = J[k * N:(k + 1) * N, z], ref[k * N:(k + 1) * N]
                e1.append(np.mean(np.abs(a - b)) / np.mean(np.abs(b)))
                emax.append(np.max(np.abs(a - b)) / np.max(np.abs(b)))
        l1[dt] = max(e1)
        say(f"  Savannah dt {dt}: jvp vs ForwardDiff L1 rel {max(e1):.2e}, max rel {max(emax):.2e}")
        assert max(emax) < 1e-2
    assert l1[0.02] < 1e-3 and l1[0.02] < 0.6 * l1[0.05]
    # the adjoint of w . Q(T) gives the same zone gradient as the tangents
    w = np.random.default_rng(0).normal(size=3 * N)
    _, nb, _, _ = ctx.adjoint(c.Q0, 0.02, int(round(T / 0.02)), 50, w, params=zn)
    gz = np.bincount(c.zone, weights=nb, minlength=len(zn))
    assert np.allclose(gz, w @ J, rtol=1e-9, atol=1e-12 * np.max(np.abs(gz)))
    say(f"  adjoint zone gradient vs fixture w.S: {np.max(np.abs(gz - w @ S)) / np.max(np.abs(w @ S)):.2e}")
