"""NumPy statement of the UDE closure n = 0.01 + softplus(MLP(h, |u|, zb))
used to check libhgswe's matrix-core evaluation (test code only)."""
import numpy as np

EPS = np.finfo(np.float64).eps


def features(case, Q):
    N = case.N
    xi, qx, qy = Q[:N], Q[N:2 * N].copy(), Q[2 * N:].copy()
    h = xi + case.hstill
    dry = ~(h > case.h_small)
    h = np.where(dry, case.h_small, h)
    qx[dry] = 0.0
    qy[dry] = 0.0
    u, v = qx / h, qy / h
    return np.stack([h, np.sqrt(u * u + v * v + EPS), case.zb], axis=1)


def mlp_n(W, F, dtype=np.float64):
    W = np.asarray(W, dtype=dtype)
    F = np.asarray(F, dtype=dtype)
    W1, b1 = W[0:48].reshape(16, 3), W[48:64]
    W2, b2 = W[64:320].reshape(16, 16), W[320:336]
    W3, b3 = W[336:352], W[352]
    a1 = np.tanh(F @ W1.T + b1)
    a2 = np.tanh(a1 @ W2.T + b2)
    z3 = a2 @ W3 + b3
    sp = np.where(z3 > 20, z3, np.log1p(np.exp(np.minimum(z3, 20))))
    return 0.01 + sp.astype(np.float64)
