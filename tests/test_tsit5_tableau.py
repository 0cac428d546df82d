"""The Tsit5 tableau libhgswe integrates with (read through the C ABI, no
GPU needed) satisfies the Runge-Kutta order conditions: row sums equal the
nodes, the weights are fifth order (all 17 conditions), the embedded
weights fourth order but not fifth, and the error weights sum to zero."""
import numpy as np

from hgswe import capi


def _tableau():
    c, a6, bt = capi.tsit5_tableau()
    A = np.zeros((7, 7))
    A[:, :6] = a6
    return c, A, A[6].copy(), bt


def _residuals(w, A, c, order):
    r = [w.sum() - 1.0]
    if order >= 2:
        r.append(w @ c - 1 / 2)
    if order >= 3:
        r += [w @ c**2 - 1 / 3, w @ A @ c - 1 / 6]
    if order >= 4:
        r += [w @ c**3 - 1 / 4, w @ (c * (A @ c)) - 1 / 8, w @ A @ c**2 - 1 / 12, w @ A @ A @ c - 1 / 24]
    if order >= 5:
        r += [w @ c**4 - 1 / 5, w @ (c**2 * (A @ c)) - 1 / 10, w @ ((A @ c) ** 2) - 1 / 20,
              w @ (c * (A @ c**2)) - 1 / 15, w @ A @ c**3 - 1 / 20, w @ (c * (A @ A @ c)) - 1 / 30,
              w @ A @ (c * (A @ c)) - 1 / 40, w @ A @ A @ c**2 - 1 / 60, w @ A @ A @ A @ c - 1 / 120]
    return np.abs(np.array(r))


def test_row_sums_are_nodes():
    c, A, _, _ = _tableau()
    assert np.max(np.abs(A.sum(axis=1) - c)) < 1e-14
    assert c[0] == 0.0 and c[6] == 1.0 and A[6, 6] == 0.0


def test_fifth_order_weights():
    c, A, b, _ = _tableau()
    assert np.max(_residuals(b, A, c, 5)) < 1e-14


def test_embedded_fourth_order_weights():
    c, A, b, bt = _tableau()
    bh = b - bt
    assert np.max(_residuals(bh, A, c, 4)) < 1e-14
    assert np.max(_residuals(bh, A, c, 5)) > 1e-4  # a genuine lower-order pair
    assert abs(bt.sum()) < 1e-15
