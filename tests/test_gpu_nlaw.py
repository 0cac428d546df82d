"""GPU checks of the variable Manning-n laws (hg_set_n_law), the forward
simulation's "variable" ManningN option.

* The device laws against the reference's own fixtures: n (and h/ks, f, Re)
  of the sigmoid n(h) bump channel and of the Savannah River h_Umag_ks case,
  evaluated at their final states (1e-14 relative; the device pow / exp /
  log10 may differ from Julia's in the last bit).
* All four laws against the oracle's NumPy statement on random inputs.
* Inside the RHS: the law's n from the state against the oracle's features
  + law; a step with the law on is bitwise the per-cell-n step given that n;
  N steps match the oracle driven by the NumPy law (1e-10).
* The reference's final states: the RHS with the law vanishes at the n(h)
  bump's steady state (and does not with constant n); the device Tsit5 from
  the reference's initial conditions reaches both fixtures' final states
  within the fixtures' own integration tolerance (reltol 1e-3).
* Derivatives through the law: one RHS cotangent against the oracle's Dual
  tangent plus central differences of the smooth NumPy law; over many steps
  the device tangent and adjoint satisfy w.(J v) = (J^T w).v.  (Central
  differences of whole trajectories are not a pin here: the reference's
  eps-smoothed |u_n| is a kink where a face's normal velocity crosses zero,
  and such differences wander at 1e-5 with constant n as well.)
"""
import json
import os
import sys

import numpy as np
import pytest

import nlaw
import pyoracle
from conftest import GOLDEN, load_golden_case
from hgswe import capi, cases

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _options(name):
    with open(os.path.join(GOLDEN, name, "run_control.json")) as fh:
        return json.load(fh)


def _truth(name):
    return np.load(os.path.join(GOLDEN, name, "truth.npz"))


def bump_nh():
    """The n(h) bump channel: case, law parameters, end time."""
    rc = _options("bump_nh")
    o = rc["forward_simulation_options"]
    ic = tuple(o["forward_simulation_initial_condition_constant_values"])
    c = load_golden_case("bump", ic=ic)
    p = {k: float(v) for k, v in o["forward_simulation_ManningN_function_parameters"].items()}
    return c, dict(kind=nlaw.SIGMOID, **p), float(rc["time_settings"]["tspan"][1])


def savannah_ks(tmp_path):
    """The Savannah River h_Umag_ks case: case, law parameters, end time."""
    from test_oracle_golden import savannah_ic_path

    rc = _options("savannah_ks")
    c = load_golden_case("savannah", ic=savannah_ic_path(tmp_path))
    ks = np.asarray(rc["forward_simulation_options"]["forward_simulation_ManningN_function_parameters"]["ks"],
                    dtype=np.float64)[c.zone]
    return c, dict(kind=nlaw.H_UMAG_KS, ks=ks), float(rc["time_settings"]["tspan"][1])


def set_law(ctx, law):
    ctx.set_n_law(law["kind"], n_lower=law.get("n_lower", 0.0), n_upper=law.get("n_upper", 0.0),
                  k=law.get("k", 1.0), h_mid=law.get("h_mid", 1.0), ks=law.get("ks"))


def truth_state(c, t):
    return np.concatenate([t["xi_truth"], t["u_truth"] * (t["h_truth"] + c.h_small),
                           t["v_truth"] * (t["h_truth"] + c.h_small)])


def test_law_eval_matches_reference_fixtures(tmp_path):
    for name, (c, law, _) in (("bump_nh", bump_nh()), ("savannah_ks", savannah_ks(tmp_path))):
        t = _truth(name)
        ctx = capi.Context(c)
        set_law(ctx, law)
        U = np.sqrt(t["u_truth"] ** 2 + t["v_truth"] ** 2)
        n, hks, f, Re = ctx.n_law_eval(t["h_truth"], U)
        fields = [(n, "ManningN_cells_truth")]
        if law["kind"] == nlaw.H_UMAG_KS:
            fields += [(hks, "h_ks_cells_truth"), (f, "friction_factor_cells_truth"), (Re, "Re_cells_truth")]
        else:
            assert not hks.any() and not f.any() and not Re.any()
        for got, key in fields:
            rel = float(np.max(np.abs(got - t[key]) / np.abs(t[key])))
            say(f"  {name} {key}: max rel {rel:.2e}")
            assert rel <= 1e-14


LAWS = [dict(kind=nlaw.POWER, n_lower=0.02, n_upper=0.05, k=0.7),
        dict(kind=nlaw.SIGMOID, n_lower=0.03, n_upper=0.06, k=20.0, h_mid=1.5),
        dict(kind=nlaw.INVERSE, n_lower=0.02, n_upper=0.06, k=2.0),
        dict(kind=nlaw.H_UMAG_KS)]


def flowing_case(seed=3, n=3000):
    """A small channel whose every cell moves (the h_Umag_ks law has n = inf at rest)."""
    c = cases.channel(n, seed=seed, Lx=1200.0, Ly=120.0, Q_in=40.0, nsteps=30, n_bumps=5)
    N = c.N
    h = c.Q0[:N] + c.hstill
    Q0 = c.Q0.copy()
    Q0[N:2 * N] = 0.3 * h
    Q0[2 * N:] = 0.02 * h
    rng = np.random.default_rng(seed)
    ks = rng.uniform(0.02, 0.3, N)
    return c, Q0, ks


def _law(law, ks):
    return dict(law, ks=ks) if law["kind"] == nlaw.H_UMAG_KS else law


@pytest.mark.parametrize("li", range(len(LAWS)))
def test_law_matches_oracle_on_random_inputs(li):
    c, _, ks = flowing_case()
    law = _law(LAWS[li], ks)
    rng = np.random.default_rng(li)
    h = rng.uniform(0.01, 5.0, c.N)
    U = rng.uniform(0.01, 3.0, c.N)
    ctx = capi.Context(c)
    set_law(ctx, law)
    got = ctx.n_law_eval(h, U)
    kw = {k: v for k, v in law.items() if k not in ("kind", "ks")}
    ref = nlaw.evaluate(law["kind"], h, U, law.get("ks"), **kw)
    for g, r in zip(got, ref):
        scale = np.where(r == 0.0, 1.0, np.abs(r))
        rel = float(np.max(np.abs(g - r) / scale))
        assert rel <= 1e-13, rel


@pytest.mark.parametrize("li", range(len(LAWS)))
def test_law_in_rhs_and_steps(li):
    c, Q0, ks = flowing_case()
    law = _law(LAWS[li], ks)
    kw = {k: v for k, v in law.items() if k not in ("kind", "ks")}
    ctx = capi.Context(c)
    set_law(ctx, law)
    n = ctx.cell_n(Q0)
    ref = nlaw.rhs_n(law["kind"], Q0, c.hstill, c.h_small, law.get("ks"), **kw)
    assert float(np.max(np.abs(n - ref) / ref)) <= 1e-13
    # a step with the law is the per-cell-n step given the law's n
    plain = capi.Context(c.copy_fields(n=n))
    assert np.array_equal(ctx.step(Q0, c.dt, 1), plain.step(Q0, c.dt, 1))
    assert np.array_equal(ctx.rhs(Q0), plain.rhs(Q0))
    # steps against the oracle driven by the NumPy law
    nsteps = 40
    qg = ctx.step(Q0, c.dt, nsteps)
    q = Q0.copy()
    orc = pyoracle.Oracle(c)
    for _ in range(nsteps):
        orc.set_fields(n=nlaw.rhs_n(law["kind"], q, c.hstill, c.h_small, law.get("ks"), **kw))
        q = orc.run(q, c.dt, 1)
    rel = float(np.max(np.abs(qg - q)) / np.max(np.abs(q)))
    say(f"  law {law['kind']}: {nsteps} steps vs oracle + NumPy law: rel {rel:.2e}")
    assert rel <= 1e-10
    # turning the law off restores the field n
    ctx.set_n_law(None)
    assert np.array_equal(ctx.step(Q0, c.dt, 3), capi.Context(c).step(Q0, c.dt, 3))


def test_rhs_vanishes_at_reference_n_of_h_steady_state():
    c, law, _ = bump_nh()
    t = _truth("bump_nh")
    Qt = truth_state(c, t)
    up = c.mesh.centroid[:, 0] < 10.0  # subcritical, steady reach upstream of the crest
    scale = 0.5 * c.g * np.max(t["h_truth"]) ** 2 / np.sqrt(np.min(c.mesh.area))
    ctx = capi.Context(c)
    r_const = np.max(np.abs(ctx.rhs(Qt).reshape(3, c.N)[:, up])) / scale
    set_law(ctx, law)
    r_law = np.max(np.abs(ctx.rhs(Qt).reshape(3, c.N)[:, up])) / scale
    say(f"  n(h) bump steady residual upstream: law {r_law:.2e}, constant n {r_const:.2e}")
    assert r_law < 1e-6 and r_const > 1e-3


def test_tsit5_with_law_reaches_reference_final_states(tmp_path):
    for name, (c, law, T) in (("bump_nh", bump_nh()), ("savannah_ks", savannah_ks(tmp_path))):
        t = _truth(name)
        ctx = capi.Context(c)
        set_law(ctx, law)
        Q, st = ctx.tsit5(c.Q0, 0.0, T, 1, abstol=1e-8, reltol=1e-6, dt0=1e-3)
        N = c.N
        h = Q[-1, :N] + c.hstill
        u = Q[-1, N:2 * N] / (h + c.h_small)
        eh = np.mean(np.abs(h - t["h_truth"])) / np.mean(t["h_truth"])
        eu = np.mean(np.abs(u - t["u_truth"])) / np.mean(np.abs(t["u_truth"]))
        # the same run with the constant field n lands elsewhere
        Qc, _ = capi.Context(c).tsit5(c.Q0, 0.0, T, 1, abstol=1e-8, reltol=1e-6, dt0=1e-3)
        hc = Qc[-1, :N] + c.hstill
        ec = np.mean(np.abs(hc - t["h_truth"])) / np.mean(t["h_truth"])
        say(f"  {name}: Tsit5 steps {st}; final h L1 {eh:.2e}, u L1 {eu:.2e}; constant n: h L1 {ec:.2e}")
        assert eh < 1e-3 and eu < 5e-3
        assert ec > 5 * eh


@pytest.mark.parametrize("li", range(len(LAWS)))
def test_rhs_cotangent_through_law(li):
    """One RHS: the device cotangent against an independent tangent, the
    oracle's Dual JVP at fixed n plus the law's own tangent by central
    differences of the NumPy law (a smooth function of the state)."""
    c, Q0, ks = flowing_case()
    law = _law(LAWS[li], ks)
    kw = {k: v for k, v in law.items() if k not in ("kind", "ks")}
    ctx = capi.Context(c)
    set_law(ctx, law)
    rng = np.random.default_rng(20 + li)
    v = np.concatenate([rng.normal(size=c.N) * 1e-2, rng.normal(size=2 * c.N) * 1e-3])
    w = rng.normal(size=3 * c.N)

    def n_of(q):
        return nlaw.rhs_n(law["kind"], q, c.hstill, c.h_small, law.get("ks"), **kw)

    e = 1e-5
    dn = (n_of(Q0 + e * v) - n_of(Q0 - e * v)) / (2.0 * e)
    orc = pyoracle.Oracle(c)
    orc.set_fields(n=n_of(Q0))
    _, Jv = orc.rhs_jvp(Q0, dQ=v, dn=dn)
    Qb = ctx.rhs_vjp(Q0, w)[0]
    a, b = float(w @ Jv), float(Qb @ v)
    say(f"  law {law['kind']}: w.Jv {a:.12e}  J^T w.v {b:.12e}")
    assert abs(a - b) <= 1e-8 * abs(a)


@pytest.mark.parametrize("li", [1, 3])
def test_steps_forward_and_reverse_agree_through_law(li):
    """nsteps Euler steps: the device tangent and the device adjoint give
    w.(J v) = (J^T w).v, and the tangent's primal is the plain step."""
    c, Q0, ks = flowing_case()
    law = _law(LAWS[li], ks)
    ctx = capi.Context(c)
    set_law(ctx, law)
    nsteps, kck = 20, 6
    rng = np.random.default_rng(11)
    v = np.concatenate([rng.normal(size=c.N) * 1e-2, rng.normal(size=2 * c.N) * 1e-3])
    w = rng.normal(size=3 * c.N)
    QN, dQN = ctx.jvp(Q0, dQ0=v, dt=c.dt, nsteps=nsteps)
    assert np.array_equal(QN, ctx.step(Q0, c.dt, nsteps))
    Q0b, _, _, _ = ctx.adjoint(Q0, c.dt, nsteps, kck, w)
    a, b = float(w @ dQN), float(Q0b @ v)
    say(f"  law {law['kind']}: w.Jv {a:.12e}  J^T w.v {b:.12e}")
    assert abs(a - b) <= 1e-10 * abs(a)


def test_law_errors():
    c, Q0, ks = flowing_case()
    ctx = capi.Context(c)
    with pytest.raises(capi.HGError):
        ctx.set_n_law("power_law", n_lower=0.02, n_upper=0.05, k=0.0)
    with pytest.raises(capi.HGError):
        ctx.set_n_law("sigmoid", n_lower=0.02, n_upper=0.05, k=1.0, h_mid=0.0)
    with pytest.raises(capi.HGError):
        ctx.set_n_law("h_Umag_ks")  # no ks
    ctx.set_n_law("inverse", n_lower=0.02, n_upper=0.05, k=1.0)
    with pytest.raises(capi.HGError):
        ctx.set_active_param(capi.HG_PARAM_N_CELL)
    ctx.set_n_law(None)
    ctx.set_active_param(capi.HG_PARAM_N_CELL)
    with pytest.raises(capi.HGError):
        ctx.set_n_law("inverse", n_lower=0.02, n_upper=0.05, k=1.0)
