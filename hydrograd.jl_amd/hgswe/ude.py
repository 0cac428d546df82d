"""Weights of the UDE closure for Manning's n (hg_set_ude): a 3 -> 16 -> 16
-> 1 tanh network, 353 values per set laid out as W1[16][3], b1[16],
W2[16][16], b2[16], W3[16], b3."""
from __future__ import annotations

import numpy as np

N_WEIGHTS = 353
SLICES = {"W1": (0, 48), "b1": (48, 64), "W2": (64, 320), "b2": (320, 336), "W3": (336, 352), "b3": (352, 353)}


def xavier_weights(seed=0, b3=-4.0):
    """Glorot-uniform weights, zero hidden biases; the output bias b3 sets
    the initial n near 0.01 + softplus(b3) (about 0.028 for b3 = -4)."""
    rng = np.random.default_rng(seed)
    w = np.zeros(N_WEIGHTS)

    def glorot(fan_in, fan_out):
        a = np.sqrt(6.0 / (fan_in + fan_out))
        return rng.uniform(-a, a, fan_in * fan_out)

    w[0:48] = glorot(3, 16)
    w[64:320] = glorot(16, 16)
    w[336:352] = glorot(16, 1)
    w[352] = b3
    return w


def perturbed_sets(base, n_sets, scale=0.05, seed=1):
    """n_sets weight sets: the base and base + scale * N(0, 1) perturbations
    (batched ensemble / finite-difference evaluations)."""
    rng = np.random.default_rng(seed)
    out = np.repeat(base[None, :], n_sets, axis=0)
    out[1:] += scale * rng.normal(size=(n_sets - 1, N_WEIGHTS))
    return out
