"""GPU checks of the reference's UDE network on the device (hg_set_lux_ude):
the three UDE choices (ManningN_h, ManningN_h_Umag_ks, FlowResistance).

* The network's output per cell against the oracle's NumPy statement
  (1e-13 relative), for both register widths of the kernels.
* ManningN choices: a step with the network on is bitwise the per-cell-n
  step given its n.  FlowResistance: the RHS against the oracle with the
  network's friction magnitudes (m q / |q|_eps friction); 40 steps against
  the oracle driven by the NumPy network (1e-10).
* Reverse mode: the state cotangent and the weight cotangent of one RHS
  against central differences of the oracle RHS with the NumPy network (a
  smooth function at these states); the weight gradient of a multi-step
  probe loss against central differences of device runs.
* The weights as the active parameter (HG_PARAM_UDE, the reference's
  params_vector) and the exclusions.
* Forward mode through the network (hg_jvp, the reference's AutoForwardDiff
  UDE path): the primal run is bitwise the stepping run; along the weights
  and along the initial state, <ybar, J v> equals <J^T ybar, v> from the
  reverse sweep (hg_adjoint with the weight cotangent); along the weights,
  against fourth-order central differences of device runs.
"""
import sys

import numpy as np
import pytest

import lux_ude as lx
import pyoracle
from hgswe import capi, cases

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def flowing_case(seed=4, n=3000):
    c = cases.channel(n, seed=seed, Lx=1200.0, Ly=120.0, Q_in=40.0, nsteps=30, n_bumps=5)
    N = c.N
    h = c.Q0[:N] + c.hstill
    Q0 = c.Q0.copy()
    Q0[N:2 * N] = 0.3 * h
    Q0[2 * N:] = 0.02 * h
    ks = np.random.default_rng(seed).uniform(0.02, 0.3, N)
    return c, Q0, ks


NETS = {
    "h": lx.Net(lx.MANNING_H, [3, 3], [lx.TANH, lx.TANH], (0.02, 0.06), (1.0, 4.0)),
    "h_umag_ks": lx.Net(lx.MANNING_H_UMAG_KS, [3, 3], [lx.TANH, lx.TANH], (0.02, 0.04), (1.0, 4.0), (0.1, 1.2),
                        (0.02, 0.3)),
    "flow_resistance": lx.Net(lx.FLOW_RESISTANCE, [3, 3], [lx.TANH, lx.TANH], (0.0, 0.02), (1.0, 4.0), (0.0, 1.0)),
    "wide": lx.Net(lx.MANNING_H_UMAG_KS, [12, 9, 5], [lx.SOFTPLUS, lx.LEAKYRELU, lx.SIGMOID], (0.02, 0.05),
                   (1.0, 4.0), (0.1, 1.2), (0.02, 0.3)),
}


def device_net(net):
    return capi.lux_net(net.choice, net.hidden, net.acts, net.output_bounds, net.h_bounds, net.u_bounds,
                        net.ks_bounds)


def make(c, net, P, ks):
    ctx = capi.Context(c)
    ctx.set_lux_ude(device_net(net), P, ks=ks if net.choice == lx.MANNING_H_UMAG_KS else None)
    return ctx


def ref_out(c, net, P, Q, ks):
    return lx.rhs_output(net, P, Q, c.hstill, c.h_small, ks)


@pytest.mark.parametrize("name", list(NETS))
def test_network_matches_oracle(name):
    c, Q0, ks = flowing_case()
    net = NETS[name]
    P = lx.random_params(net, seed=1)
    assert capi.lux_n_params(device_net(net)) == net.n_params
    ctx = make(c, net, P, ks)
    Q = Q0.copy()
    Q[: c.N] += np.random.default_rng(2).normal(0.0, 0.3, c.N)  # spread the depths over the bounds
    got = ctx.lux_eval(Q)
    ref = ref_out(c, net, P, Q, ks)
    rel = float(np.max(np.abs(got - ref) / np.abs(ref)))
    say(f"  {name}: network output max rel {rel:.2e}, range [{ref.min():.4g}, {ref.max():.4g}]")
    assert rel <= 1e-13


@pytest.mark.parametrize("name", ["h", "h_umag_ks", "wide"])
def test_manning_network_step_is_per_cell_n_step(name):
    c, Q0, ks = flowing_case()
    net = NETS[name]
    P = lx.random_params(net, seed=2)
    ctx = make(c, net, P, ks)
    n = ctx.lux_eval(Q0)
    plain = capi.Context(c.copy_fields(n=n))
    assert np.array_equal(ctx.step(Q0, c.dt, 1), plain.step(Q0, c.dt, 1))
    assert np.array_equal(ctx.rhs(Q0), plain.rhs(Q0))
    assert np.array_equal(ctx.cell_n(Q0), n)


@pytest.mark.parametrize("name", ["h", "flow_resistance"])
def test_network_steps_match_oracle(name):
    c, Q0, ks = flowing_case()
    net = NETS[name]
    P = lx.random_params(net, seed=3)
    ctx = make(c, net, P, ks)
    orc = pyoracle.Oracle(c)

    def orc_set(q):
        out = ref_out(c, net, P, q, ks)
        if net.choice == lx.FLOW_RESISTANCE:
            orc.set_fields(fmag=out)
        else:
            orc.set_fields(n=out)

    orc_set(Q0)
    r_dev, r_orc = ctx.rhs(Q0), orc.rhs(Q0)
    rel = float(np.max(np.abs(r_dev - r_orc)) / np.max(np.abs(r_orc)))
    say(f"  {name}: RHS vs oracle rel {rel:.2e}")
    assert rel <= 1e-13
    nsteps = 40
    qg = ctx.step(Q0, c.dt, nsteps)
    q = Q0.copy()
    for _ in range(nsteps):
        orc_set(q)
        q = orc.run(q, c.dt, 1)
    rel = float(np.max(np.abs(qg - q)) / np.max(np.abs(q)))
    say(f"  {name}: {nsteps} steps vs oracle + NumPy network: rel {rel:.2e}")
    assert rel <= 1e-10


def _orc_rhs(c, net, P, q, ks):
    o = pyoracle.Oracle(c)
    out = ref_out(c, net, P, q, ks)
    o.set_fields(**({"fmag": out} if net.choice == lx.FLOW_RESISTANCE else {"n": out}))
    return o.rhs(q)


@pytest.mark.parametrize("name", ["h", "h_umag_ks", "flow_resistance", "wide"])
def test_rhs_cotangents_against_differences(name):
    c, Q0, ks = flowing_case()
    net = NETS[name]
    P = lx.random_params(net, seed=5)
    ctx = make(c, net, P, ks)
    rng = np.random.default_rng(6)
    w = rng.normal(size=3 * c.N)
    Qb = ctx.rhs_vjp(Q0, w)[0]
    Pb = ctx.lux_grad()
    # state direction
    v = np.concatenate([rng.normal(size=c.N) * 1e-2, rng.normal(size=2 * c.N) * 1e-3])
    e = 1e-4
    fd = float(w @ (_orc_rhs(c, net, P, Q0 + e * v, ks) - _orc_rhs(c, net, P, Q0 - e * v, ks))) / (2 * e)
    b = float(Qb @ v)
    say(f"  {name}: state dir fd {fd:.10e} adjoint {b:.10e}")
    assert abs(fd - b) <= 1e-6 * abs(b)
    # weight direction (the step stays below the distance of the nearest
    # leakyrelu pre-activation from its kink, about 6e-6 for these weights)
    u = rng.normal(size=P.size)
    e = 1e-6
    fd = float(w @ (_orc_rhs(c, net, P + e * u, Q0, ks) - _orc_rhs(c, net, P - e * u, Q0, ks))) / (2 * e)
    b = float(Pb @ u)
    say(f"  {name}: weight dir fd {fd:.10e} adjoint {b:.10e}")
    assert abs(fd - b) <= 1e-6 * abs(b)


@pytest.mark.parametrize("name", ["h_umag_ks", "flow_resistance"])
def test_multistep_weight_gradient(name):
    c, Q0, ks = flowing_case()
    net = NETS[name]
    P = lx.random_params(net, seed=7)
    nsteps, kck = 20, 6
    rng = np.random.default_rng(8)
    probe = rng.choice(c.N, size=25, replace=False).astype(np.int32)
    ctx = make(c, net, P, ks)
    QN = ctx.step(Q0, c.dt, nsteps)
    target = QN[probe] + c.hstill[probe] + c.zb[probe] + rng.normal(0.0, 0.02, probe.size)
    ctx.grad_probes(Q0, c.dt, nsteps, kck, probe, target)
    g = ctx.lux_grad()
    again = make(c, net, P, ks)
    again.grad_probes(Q0, c.dt, nsteps, kck, probe, target)
    assert np.array_equal(g, again.lux_grad())  # deterministic

    def L(p):
        k = make(c, net, p, ks)
        q = k.step(Q0, c.dt, nsteps)
        r = q[probe] + c.hstill[probe] + c.zb[probe] - target
        return float(np.sum(r * r))

    for trial in range(2):
        u = rng.normal(size=P.size)
        e = 1e-3
        fd = (8.0 * (L(P + e * u) - L(P - e * u)) - (L(P + 2 * e * u) - L(P - 2 * e * u))) / (12.0 * e)
        b = float(g @ u)
        say(f"  {name}: weights dir {trial}: fd {fd:.10e} adjoint {b:.10e} rel {abs(fd - b) / abs(b):.2e}")
        assert abs(fd - b) <= 1e-5 * abs(b)


def test_weights_as_active_parameter():
    c, Q0, ks = flowing_case()
    net = NETS["h_umag_ks"]
    P1 = lx.random_params(net, seed=9)
    P2 = lx.random_params(net, seed=10)
    ctx = make(c, net, P1, ks)
    ctx.set_active_param(capi.HG_PARAM_UDE)
    r = ctx.rhs(Q0, params=P2)
    assert np.array_equal(r, make(c, net, P2, ks).rhs(Q0))
    with pytest.raises(capi.HGError):
        ctx.rhs(Q0, params=P2[:-1])


def test_network_exclusions():
    c, Q0, ks = flowing_case()
    net = NETS["h"]
    P = lx.random_params(net, seed=11)
    ctx = make(c, net, P, ks)
    with pytest.raises(capi.HGError):
        ctx.set_active_param(capi.HG_PARAM_N_CELL)
    with pytest.raises(capi.HGError):
        ctx.set_n_law("inverse", n_lower=0.02, n_upper=0.05, k=1.0)
    ctx.set_lux_ude(None)
    assert np.array_equal(ctx.step(Q0, c.dt, 2), capi.Context(c).step(Q0, c.dt, 2))
    with pytest.raises(capi.HGError):
        ctx.set_active_param(capi.HG_PARAM_UDE)
    bad = lx.Net(lx.MANNING_H, [17], [lx.TANH], (0.02, 0.06), (1.0, 4.0))
    assert capi.lux_n_params(device_net(bad)) == -1


@pytest.mark.parametrize("name", ["h", "h_umag_ks", "flow_resistance"])
def test_forward_mode_through_the_network(name):
    c, Q0, ks = flowing_case()
    net = NETS[name]
    P = lx.random_params(net, seed=21)
    ctx = make(c, net, P, ks)
    ctx.set_active_param(capi.HG_PARAM_UDE)
    nsteps, kck = 12, 5
    rng = np.random.default_rng(22)
    dW = rng.normal(size=P.size)
    QN, dQN = ctx.jvp(Q0, c.dt, nsteps, params=P, dparams=dW)
    assert np.array_equal(QN, ctx.step(Q0, c.dt, nsteps, params=P))
    ybar = rng.normal(size=3 * c.N)
    Q0b, _, _, _ = ctx.adjoint(Q0, c.dt, nsteps, kck, ybar, params=P)
    gW = ctx.lux_grad()
    lhs, rhs = float(np.dot(ybar, dQN)), float(np.dot(gW, dW))
    scale = float(np.sum(np.abs(ybar * dQN)))
    say(f"  {name} JVP along the weights: <y, Jv> {lhs:.12e}, <J^T y, v> {rhs:.12e}")
    assert abs(lhs - rhs) <= 1e-10 * scale
    # along the initial state, the weights held
    dQ0 = rng.normal(size=3 * c.N) * 1e-3
    _, dQN2 = ctx.jvp(Q0, c.dt, nsteps, dQ0=dQ0, params=P)
    lhs2, rhs2 = float(np.dot(ybar, dQN2)), float(np.dot(Q0b, dQ0))
    scale2 = float(np.sum(np.abs(ybar * dQN2)))
    say(f"  {name} JVP along Q0: <y, Jv> {lhs2:.12e}, <J^T y, v> {rhs2:.12e}")
    assert abs(lhs2 - rhs2) <= 1e-10 * scale2
    # fourth-order central differences of device runs along the weights
    e = 1e-4

    def f(w):
        return float(np.dot(ybar, ctx.step(Q0, c.dt, nsteps, params=w)))

    fd = (8.0 * (f(P + e * dW) - f(P - e * dW)) - (f(P + 2 * e * dW) - f(P - 2 * e * dW))) / (12.0 * e)
    say(f"  {name} JVP along the weights vs differences: {lhs:.10e} / {fd:.10e}")
    assert abs(fd - lhs) <= 1e-6 * abs(lhs) + 1e-12

