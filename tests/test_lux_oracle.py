"""Pins of the oracle's UDE network statement (oracle/lux_ude.py).

No Lux / NNlib is in the reference tree, so the published layer semantics
are pinned two ways:
* against torch's float64 Linear, LayerNorm (biased variance) and
  activations, an independent implementation of the same definitions;
* against closed forms: a one-unit LayerNorm outputs its bias; a two-unit
  LayerNorm outputs +-d / sqrt(d^2 + eps) (the uncorrected variance); zero
  output weights give the middle of the output bounds; the input
  normalisation maps each bound to -1 / +1.
"""
import numpy as np
import pytest
import torch

import lux_ude as lx


def torch_forward(net, P, F):
    offs, out, _ = net.layout()
    x = torch.tensor(F.T, dtype=torch.float64)  # (N, n_in)
    acts = {lx.RELU: torch.relu, lx.LEAKYRELU: lambda t: torch.nn.functional.leaky_relu(t, 0.01),
            lx.SIGMOID: torch.sigmoid, lx.TANH: torch.tanh, lx.SOFTPLUS: torch.nn.functional.softplus}
    for o, k in zip(offs, net.acts):
        w, win = o["w"], o["in"]
        lin = torch.nn.Linear(win, w, dtype=torch.float64)
        ln = torch.nn.LayerNorm(w, eps=net.eps, dtype=torch.float64)
        with torch.no_grad():
            lin.weight.copy_(torch.tensor(P[o["W"]:o["W"] + w * win].reshape(win, w).T))
            lin.bias.copy_(torch.tensor(P[o["b"]:o["b"] + w]))
            ln.weight.copy_(torch.tensor(P[o["g"]:o["g"] + w]))
            ln.bias.copy_(torch.tensor(P[o["be"]:o["be"] + w]))
            x = ln(acts[k](lin(x)))
    with torch.no_grad():
        o = x @ torch.tensor(P[out["W"]:out["W"] + out["in"]]) + P[out["b"]]
        lo, hi = net.output_bounds
        return (lo + (hi - lo) * torch.sigmoid(o)).numpy()


@pytest.mark.parametrize("choice,hidden,acts", [
    (lx.MANNING_H, [3, 3], [lx.TANH, lx.TANH]),
    (lx.MANNING_H_UMAG_KS, [5, 4, 3], [lx.RELU, lx.SIGMOID, lx.LEAKYRELU]),
    (lx.FLOW_RESISTANCE, [16, 7], [lx.SOFTPLUS, lx.TANH]),
    (lx.MANNING_H_UMAG_KS, [2, 2, 2, 2], [lx.LEAKYRELU, lx.SOFTPLUS, lx.RELU, lx.SIGMOID]),
])
def test_network_matches_torch(choice, hidden, acts):
    net = lx.Net(choice, hidden, acts, (0.02, 0.05), (0.1, 5.0), (0.0, 2.0), (0.01, 0.3))
    P = lx.random_params(net, seed=len(hidden))
    rng = np.random.default_rng(1)
    F = rng.uniform(-1.5, 1.5, (net.n_in, 500))
    y = lx.forward(net, P, F)
    yt = torch_forward(net, P, F)
    assert np.max(np.abs(y - yt) / yt) < 1e-14
    assert np.all((y > 0.02) & (y < 0.05))


def test_layer_norm_closed_forms():
    # one unit: the normalised value is 0, so the layer outputs its bias and
    # the network is constant
    net = lx.Net(lx.MANNING_H, [1], [lx.TANH], (0.01, 0.05), (0.0, 1.0))
    P = np.array([0.7, 0.1, 3.0, 0.4, 2.0, -0.5])  # W, b, scale, bias, Wo, bo
    y = lx.forward(net, P, np.array([[-1.0, 0.0, 0.3, 1.0]]))
    assert np.all(y == 0.01 + 0.04 * lx.nn_sigmoid(np.array(2.0 * 0.4 - 0.5)))
    # two units with pre-activations t and -t: xhat = +-d / sqrt(d^2 + eps), d = tanh(t)
    net = lx.Net(lx.MANNING_H, [2], [lx.TANH], (0.0, 1.0), (0.0, 1.0))
    t = 0.8
    P = np.array([t, -t, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0, 1.0, 0.0, 0.0])  # W (2x1), b, scale, bias, Wo, bo
    d = np.tanh(t)
    xh = d / np.sqrt(d * d + lx.LN_EPS)
    y = lx.forward(net, P, np.array([[1.0]]))
    assert abs(y[0] - lx.nn_sigmoid(np.array(xh))) < 1e-16
    # zero output weights: the middle of the output bounds
    net = lx.Net(lx.FLOW_RESISTANCE, [3], [lx.RELU], (0.2, 0.6), (0.0, 1.0), (0.0, 1.0))
    P = lx.random_params(net, 3)
    offs, out, _ = net.layout()
    P[out["W"]:out["W"] + out["in"] + 1] = 0.0
    assert np.all(lx.forward(net, P, np.ones((2, 4))) == 0.4)


def test_input_normalisation():
    N = 4
    hstill = np.zeros(N)
    h = np.array([0.5, 2.0, 1e-4, 1.25])  # the third cell is dry: h -> h_small, q -> 0
    qx = np.array([1.0, 0.0, 3.0, 0.0])
    qy = np.array([0.0, 4.0, 1.0, 0.0])
    Q = np.concatenate([h, qx, qy])
    net = lx.Net(lx.MANNING_H_UMAG_KS, [2], [lx.TANH], (0, 1), (0.5, 2.0), (0.0, 2.0), (0.1, 0.3))
    F = lx.features(net, Q, hstill, 1e-3, ks=np.array([0.1, 0.3, 0.2, 0.1]))
    assert np.array_equal(F[0][[0, 1, 3]], [-1.0, 1.0, 0.0])
    assert np.array_equal(F[1], [1.0, 1.0, -1.0, -1.0])  # |U| = 2, 2, 0, 0
    assert np.allclose(F[2], [-1.0, 1.0, 0.0, -1.0], atol=1e-15)
    fr = lx.Net(lx.FLOW_RESISTANCE, [2], [lx.TANH], (0, 1), (0.5, 2.0), (0.0, 1.0))  # |q| bounds 0 .. 2
    F = lx.features(fr, Q, hstill, 1e-3)
    assert abs(F[1][0] - 0.0) < 1e-15 and abs(F[1][1] - 3.0) < 1e-15 and abs(F[1][2] + 1.0) < 1e-7


def test_flow_resistance_friction_closed_forms():
    """Pins of the oracle's FlowResistance friction (semi_discretize_swe_2D.jl:
    521-531: f = m q / sqrt(q_x^2 + q_y^2 + eps), masked on dry cells by the
    source term) through what the mathematics fixes, from D = RHS(m) - RHS(0),
    the friction-only change of the RHS:
    * the mass row does not change (exactly);
    * D opposes q (antiparallel, D . q < 0);
    * |D| = m for |q| >> sqrt(eps), and |D| = m |q| / sqrt(eps) for |q| << sqrt(eps)
      (the eps sits inside the square root);
    * q = 0 or a dry cell gives D = 0 exactly; D is linear in m.
    A dropped or misplaced eps, a wrong sign, swapped components or a missing
    dry mask each fails one of these."""
    import pyoracle
    from hgswe import cases

    c = cases.dam_break(nx=40, ny=4)
    N = c.N
    rng = np.random.default_rng(21)
    h = np.full(N, 0.8)
    h[:4] = c.h_small * 0.5  # dry cells (h <= h_small)
    mag = np.empty(N)
    mag[4:14] = 0.0  # zero discharge
    mag[14:24] = 1e-11  # |q| << sqrt(eps) ~ 1.5e-8
    mag[24:] = rng.uniform(0.05, 2.0, N - 24)
    mag[:4] = 0.3
    ang = rng.uniform(0.0, 2.0 * np.pi, N)
    qx, qy = mag * np.cos(ang), mag * np.sin(ang)
    Q = np.concatenate([h - c.hstill, qx, qy])
    m = rng.uniform(0.01, 0.1, N)
    orc = pyoracle.Oracle(c)

    def rhs(fm):
        orc.set_fields(fmag=fm)
        return orc.rhs(Q)

    r0, r1, r2 = rhs(np.zeros(N)), rhs(m), rhs(2.0 * m)
    D, D2 = r1 - r0, r2 - r0
    Dx, Dy = D[N:2 * N], D[2 * N:]
    assert np.array_equal(D[:N], np.zeros(N))
    dry, zero, tiny, big = slice(0, 4), slice(4, 14), slice(14, 24), slice(24, N)
    assert np.all(Dx[dry] == 0.0) and np.all(Dy[dry] == 0.0)
    assert np.all(Dx[zero] == 0.0) and np.all(Dy[zero] == 0.0)
    # large |q|: magnitude m, antiparallel to q
    Dm = np.hypot(Dx[big], Dy[big])
    assert np.max(np.abs(Dm / m[big] - 1.0)) < 1e-9
    assert np.all(Dx[big] * qx[big] + Dy[big] * qy[big] < 0.0)
    cross = np.abs(Dx[big] * qy[big] - Dy[big] * qx[big]) / (Dm * mag[big])
    assert np.max(cross) < 1e-9
    # small |q|: the linear regime of the regularised magnitude
    Dt = np.hypot(Dx[tiny], Dy[tiny])
    ratio = Dt / (m[tiny] * mag[tiny] / np.sqrt(lx.EPS))
    assert np.max(np.abs(ratio - 1.0)) < 1e-5
    # linear in m
    assert np.allclose(D2[N:], 2.0 * D[N:], rtol=1e-9, atol=1e-15)
