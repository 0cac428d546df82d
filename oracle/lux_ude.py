"""TEST INFRASTRUCTURE ONLY: NumPy restatement of Hydrograd.jl's UDE network
and of its inputs, used as the checker for the library's device network.
Only the test-suite may import this module; the product path never does.

Network (create_NN_model, src/UDE/process_UDE.jl:2-65): per hidden layer
Dense(in, w, act) -> LayerNorm(w); then Dense(w, 1) and the output bounds
lo + (hi - lo) sigmoid(x).  Lux 1.2.3 and NNlib are not in the reference
tree; their published semantics are restated:
  Dense      y = act(W x + b), W of size (out, in), stored column-major;
  LayerNorm  over the units of one sample, uncorrected variance,
             y = (a - mean) / sqrt(var + eps) * scale + bias, eps = 1f-5;
  NNlib      relu, leakyrelu (0.01), sigmoid (exp(-|x|) form), tanh,
             softplus = log1p(exp(-|x|)) + relu(x).
Parity with Lux's own arithmetic (BLAS summation order, its LayerNorm
kernel) is unpinned; tests/test_lux_oracle.py pins this statement against
torch's float64 Linear / LayerNorm / activations and against closed forms.

Inputs (normalised to [-1, 1] by bounds, 2 (x - lo) / (hi - lo) - 1):
  ManningN_h          h                      (process_ManningN_2D.jl:234)
  ManningN_h_Umag_ks  h, |U| = |q| / h, ks   (:234-239; Umag as in
                                              semi_discretize_swe_2D.jl:145-148)
  FlowResistance      h, |q| = sqrt(q_x^2 + q_y^2 + eps), |q| bounds
                      h_lo U_lo .. h_hi U_hi (process_FlowResistance_UDE.jl:14-21)
with h the RHS's clamped depth and q its masked discharge.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

EPS = np.finfo(np.float64).eps
LN_EPS = float(np.float32(1e-5))  # Lux's LayerNorm default epsilon, as Float64

MANNING_H, MANNING_H_UMAG_KS, FLOW_RESISTANCE = 1, 2, 3
N_IN = {MANNING_H: 1, MANNING_H_UMAG_KS: 3, FLOW_RESISTANCE: 2}
RELU, LEAKYRELU, SIGMOID, TANH, SOFTPLUS = range(5)


def nn_sigmoid(x):
    t = np.exp(-np.abs(x))
    return np.where(x >= 0, 1.0 / (1.0 + t), t / (1.0 + t))


def act(k, x):
    if k == RELU:
        return np.where(x < 0, 0.0, x)
    if k == LEAKYRELU:
        return np.where(x > 0, x, 0.01 * x)
    if k == SIGMOID:
        return nn_sigmoid(x)
    if k == TANH:
        return np.tanh(x)
    return np.log1p(np.exp(-np.abs(x))) + np.where(x < 0, 0.0, x)


@dataclass
class Net:
    choice: int
    hidden: list
    acts: list
    output_bounds: tuple
    h_bounds: tuple
    u_bounds: tuple = (0.0, 1.0)
    ks_bounds: tuple = (0.0, 1.0)
    eps: float = LN_EPS
    offsets: list = field(default_factory=list)

    @property
    def n_in(self):
        return N_IN[self.choice]

    def layout(self):
        """Flat weight order: per hidden layer W (column-major), b, scale, bias; then Wo, bo."""
        offs, cur, win = [], 0, self.n_in
        for w in self.hidden:
            o = {"W": cur, "b": cur + w * win, "g": cur + w * win + w, "be": cur + w * win + 2 * w, "w": w, "in": win}
            offs.append(o)
            cur += w * win + 3 * w
            win = w
        out = {"W": cur, "b": cur + win, "in": win}
        return offs, out, cur + win + 1

    @property
    def n_params(self):
        return self.layout()[2]


def forward(net: Net, P, F):
    """The chain on features F (n_in x N) -> outputs (N)."""
    P = np.asarray(P, dtype=np.float64)
    offs, out, _ = net.layout()
    x = np.asarray(F, dtype=np.float64)
    for o, k in zip(offs, net.acts):
        w, win = o["w"], o["in"]
        W = P[o["W"]:o["W"] + w * win].reshape(win, w).T  # column-major (w x win)
        z = W[:, 0:1] * x[0:1]
        for j in range(1, win):
            z = z + W[:, j:j + 1] * x[j:j + 1]
        a = act(k, z + P[o["b"]:o["b"] + w, None])
        mu = a.mean(axis=0, keepdims=True)
        var = ((a - mu) ** 2).mean(axis=0, keepdims=True)
        x = (a - mu) / np.sqrt(var + net.eps) * P[o["g"]:o["g"] + w, None] + P[o["be"]:o["be"] + w, None]
    Wo = P[out["W"]:out["W"] + out["in"]]
    o = Wo[0] * x[0]
    for j in range(1, out["in"]):
        o = o + Wo[j] * x[j]
    o = o + P[out["b"]]
    lo, hi = net.output_bounds
    return lo + (hi - lo) * nn_sigmoid(o)


def _norm(x, b):
    return 2.0 * (x - b[0]) / (b[1] - b[0]) - 1.0


def features(net: Net, Q, hstill, h_small, ks=None):
    """Normalised inputs (n_in x N) of the RHS state Q."""
    N = hstill.size
    xi, qx, qy = Q[:N], Q[N:2 * N], Q[2 * N:]
    h = xi + hstill
    dry = h <= h_small
    h = np.where(dry, h_small, h)
    qx = np.where(dry, 0.0, qx)
    qy = np.where(dry, 0.0, qy)
    F = [_norm(h, net.h_bounds)]
    if net.choice == MANNING_H_UMAG_KS:
        u, v = qx / h, qy / h
        F.append(_norm(np.sqrt(u * u + v * v), net.u_bounds))
        F.append(_norm(np.asarray(ks, dtype=np.float64), net.ks_bounds))
    elif net.choice == FLOW_RESISTANCE:
        qb = (net.h_bounds[0] * net.u_bounds[0], net.h_bounds[1] * net.u_bounds[1])
        F.append(_norm(np.sqrt(qx * qx + qy * qy + EPS), qb))
    return np.array(F)


def rhs_output(net: Net, P, Q, hstill, h_small, ks=None):
    """The network's output per cell in one RHS evaluation at state Q."""
    return forward(net, P, features(net, Q, hstill, h_small, ks))


def random_params(net: Net, seed, scale=0.5):
    """Weights for tests: Dense weights ~ U(-scale, scale), biases small, LayerNorm
    scale around 1 and bias around 0."""
    rng = np.random.default_rng(seed)
    offs, out, P = net.layout()
    p = np.zeros(P)
    for o in offs:
        w, win = o["w"], o["in"]
        p[o["W"]:o["W"] + w * win] = rng.uniform(-scale, scale, w * win) * np.sqrt(6.0 / (w + win))
        p[o["b"]:o["b"] + w] = rng.uniform(-0.1, 0.1, w)
        p[o["g"]:o["g"] + w] = 1.0 + rng.uniform(-0.2, 0.2, w)
        p[o["be"]:o["be"] + w] = rng.uniform(-0.2, 0.2, w)
    p[out["W"]:out["W"] + out["in"]] = rng.uniform(-1.0, 1.0, out["in"])
    p[out["b"]] = rng.uniform(-0.2, 0.2)
    return p
