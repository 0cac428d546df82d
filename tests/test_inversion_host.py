"""Host-side checks of the inversion driver (hgswe.inversion): the loss
terms of compute_loss_inversion and their gradients against central
differences, the Green-Gauss operator against the bed-slope data (which is
pinned bit-exactly to the reference's S0), and Adam's update rule."""
import numpy as np
import pytest

from conftest import load_golden_case
from hgswe import inversion as inv
from hgswe import mesh as hm


@pytest.fixture(scope="module")
def sav(tmp_path_factory):
    from test_oracle_golden import savannah_ic_path

    return load_golden_case("savannah", ic=savannah_ic_path(tmp_path_factory.mktemp("sav")))


def _truth(c, rng):
    N = c.N
    return {"WSE_truth": c.zb + c.hstill + rng.normal(0, 0.05, N), "u_truth": rng.normal(0, 0.3, N),
            "v_truth": rng.normal(0, 0.3, N)}


def test_gradient_operator_is_minus_bed_slope(sav):
    G = inv.gradient_operator(sav.mesh)
    _, sx, sy = hm.bed_data(sav.mesh, sav.zb)
    g = G @ sav.zb
    assert np.max(np.abs(-g[: sav.N] - sx)) < 1e-13 * max(1.0, np.max(np.abs(sx)))
    assert np.max(np.abs(-g[sav.N:] - sy)) < 1e-13 * max(1.0, np.max(np.abs(sy)))


def test_final_state_loss_gradient(sav):
    rng = np.random.default_rng(0)
    c = sav
    N = c.N
    truth = _truth(c, rng)
    QN = np.concatenate([rng.normal(0, 0.1, N), rng.normal(0, 0.5, N), rng.normal(0, 0.5, N)])
    params = rng.uniform(0.0, 0.08, 6)
    S = inv.LossSettings(wse=True, uv=True, bound=True)
    L, parts, qb, pb, zbd = inv.final_state_loss(c, QN, truth, S, params=params)
    assert parts["bound"] > 0 and parts["wse"] > 0 and parts["uv"] > 0
    v = rng.normal(size=3 * N)
    e = 1e-6
    fd = (inv.final_state_loss(c, QN + e * v, truth, S, params=params)[0]
          - inv.final_state_loss(c, QN - e * v, truth, S, params=params)[0]) / (2 * e)
    assert abs(fd - qb @ v) <= 1e-5 * abs(fd)  # FD truncation where h is small
    w = rng.normal(size=6) * 1e-3
    Sb = inv.LossSettings(wse=False, uv=False, bound=True)  # alone: the state terms would swamp it in round-off
    fd = (inv.final_state_loss(c, QN, truth, Sb, params=params + e * w)[0]
          - inv.final_state_loss(c, QN, truth, Sb, params=params - e * w)[0]) / (2 * e)
    assert abs(fd - pb @ w) <= 1e-6 * abs(fd) + 1e-15
    z = rng.normal(size=N)
    fd = (inv.final_state_loss(c, QN, truth, S, params=params, zb=c.zb + e * z)[0]
          - inv.final_state_loss(c, QN, truth, S, params=params, zb=c.zb - e * z)[0]) / (2 * e)
    assert abs(fd - zbd @ z) <= 1e-6 * abs(fd)


def test_slope_loss_gradient(sav):
    rng = np.random.default_rng(1)
    G = inv.gradient_operator(sav.mesh)
    zb = sav.zb * 3.0 + rng.normal(0, 0.5, sav.N)  # steep enough to violate the bound
    L, g = inv.slope_loss(G, zb, 0.2)
    assert L > 0
    v = rng.normal(size=sav.N)
    e = 1e-6
    fd = (inv.slope_loss(G, zb + e * v, 0.2)[0] - inv.slope_loss(G, zb - e * v, 0.2)[0]) / (2 * e)
    assert abs(fd - g @ v) <= 1e-6 * abs(fd)


def test_bound_loss_values():
    L, g = inv.bound_loss(np.array([0.005, 0.03, 0.07]), 0.01, 0.06)
    assert np.isclose(L, (0.005 / 0.05) ** 2 + (0.01 / 0.05) ** 2)
    assert g[1] == 0.0 and g[0] < 0 < g[2]


def test_adam_matches_update_rule():
    opt = inv.Adam(lr=0.1)
    p = np.array([1.0, -2.0])
    g1, g2 = np.array([0.5, -1.0]), np.array([0.25, 2.0])
    p1 = opt.step(p, g1)
    assert np.allclose(p1, p - 0.1 * np.sign(g1), atol=1e-7)  # first Adam step moves by ~lr
    m = 0.9 * 0.1 * g1 + 0.1 * g2
    v = 0.999 * 0.001 * g1 ** 2 + 0.001 * g2 ** 2
    p2 = opt.step(p1, g2)
    assert np.allclose(p2, p1 - 0.1 * (m / (1 - 0.81)) / (np.sqrt(v / (1 - 0.999 ** 2)) + 1e-8))
