"""GPU checks of the inversion driver on the Savannah River mesh (the
reference's ManningN inversion example, 6 Manning zones): the gradient of
the reference's inversion loss through the device adjoint against central
differences, a twin experiment in which Adam recovers the zone values, and
the reference's Adam -> LBFGS schedule lowering the loss past Adam's."""
import sys

import numpy as np
import pytest

from conftest import load_golden_case
from hgswe import capi, cases
from hgswe import inversion as inv

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def problem(tmp_path_factory):
    from test_oracle_golden import savannah_ic_path

    c = load_golden_case("savannah", ic=savannah_ic_path(tmp_path_factory.mktemp("sav")))
    dt = 0.2 * cases.cfl_dt(c, 1.0)
    nsteps = 150
    n_true = np.array([0.03, 0.04, 0.05, 0.03, 0.045, 0.05])
    ctx = capi.Context(c)
    ctx.set_active_param(capi.HG_PARAM_N_ZONE, zone=c.zone, n_zones=6)
    QT = ctx.step(c.Q0, dt, nsteps, params=n_true)
    N = c.N
    h = QT[:N] + c.hstill
    truth = {"WSE_truth": h + c.zb, "u_truth": QT[N:2 * N] / (h + inv.EPS), "v_truth": QT[2 * N:] / (h + inv.EPS)}
    return c, ctx, dt, nsteps, n_true, truth


def test_inversion_loss_gradient_fd(problem):
    c, ctx, dt, nsteps, n_true, truth = problem
    S = inv.LossSettings()
    p = np.array([0.03, 0.03, 0.035, 0.028, 0.04, 0.045])
    L, parts, g = inv.loss_and_grad(ctx, c, truth, p, capi.HG_PARAM_N_ZONE, dt, nsteps, 25, S, zone=c.zone)
    rng = np.random.default_rng(3)
    for trial in range(3):
        v = rng.normal(size=6) * 1e-3
        e = 1e-3
        Lp = inv.loss_and_grad(ctx, c, truth, p + e * v, capi.HG_PARAM_N_ZONE, dt, nsteps, 25, S, zone=c.zone)[0]
        Lm = inv.loss_and_grad(ctx, c, truth, p - e * v, capi.HG_PARAM_N_ZONE, dt, nsteps, 25, S, zone=c.zone)[0]
        fd = (Lp - Lm) / (2 * e)
        say(f"  inversion loss dir {trial}: fd {fd:.10e} adjoint {g @ v:.10e}")
        assert abs(fd - g @ v) <= 1e-5 * abs(fd)


def test_twin_experiment_recovers_zones(problem):
    c, ctx, dt, nsteps, n_true, truth = problem
    p0 = np.full(6, 0.03)
    res = inv.invert(ctx, c, truth, p0, capi.HG_PARAM_N_ZONE, dt, nsteps, kck=25, lr=2e-3, iters=60, zone=c.zone)
    used = np.unique(c.zone)
    err0 = np.max(np.abs(p0[used] - n_true[used]))
    err1 = np.max(np.abs(res.params[used] - n_true[used]))
    say(f"  twin: loss {res.loss[0]:.3e} -> {res.loss[-1]:.3e}; max zone error {err0:.4f} -> {err1:.4f}; "
        f"n = {np.round(res.params, 4)}")
    assert res.loss[-1] < 0.05 * res.loss[0]
    assert err1 < 0.5 * err0


def test_adam_then_lbfgs_lowers_the_loss_further(problem):
    """The examples' ["Adam", "LBFGS"] schedule (run_control.json
    inversion_optimizers; swe_2D_inversion.jl:664-729): LBFGS started from
    Adam's parameters drives the loss well below where Adam stopped."""
    c, ctx, dt, nsteps, n_true, truth = problem
    p0 = np.full(6, 0.03)
    res = inv.invert_staged(ctx, c, truth, p0, capi.HG_PARAM_N_ZONE, dt, nsteps,
                            [("Adam", 2e-3, 30), ("LBFGS", None, 15)], kck=25, zone=c.zone)
    adam = [L for k, it, L in res.stage_loss if k == 0]
    lbfgs = [L for k, it, L in res.stage_loss if k == 1]
    used = np.unique(c.zone)
    err = np.max(np.abs(res.params[used] - n_true[used]))
    say(f"  Adam -> LBFGS: loss {res.loss[0]:.3e} -> {adam[-1]:.3e} (Adam, 30 it) -> {lbfgs[-1]:.3e} "
        f"(LBFGS, {res.stages[-1][2]} it, {res.stages[-1][1]}); {len(res.loss)} evaluations; "
        f"max zone error {err:.2e}")
    assert [s[0] for s in res.stages] == ["Adam", "LBFGS"]
    assert lbfgs and lbfgs[-1] < 0.1 * adam[-1]
    assert all(b <= a for a, b in zip(lbfgs, lbfgs[1:]))  # accepted LBFGS steps never raise the loss

