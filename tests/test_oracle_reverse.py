"""The oracle's reverse mode (tape) against its forward mode (Dual).

orc_grad_run records every Euler step on an operator-overloading tape and
sweeps it backwards; orc_jvp_run / orc_jvp_state push one tangent through the
same steps.  Both differentiate the same arithmetic along the same branches,
so for any direction v, g . v from the reverse sweep equals the Dual run's
derivative of the loss up to rounding (tolerance 1e-10 relative), for the
per-cell Manning n, the inlet discharges, the initial state and the bed zb
(bed data rebuilt from zb), on a wet channel and on one with wet/dry fronts.  A step with its adjoint dropped
(the reverse sweep seeded from the wrong state) or a partial with a wrong
sign fails these by orders of magnitude; the mistake test below plants one.
"""
import numpy as np
import pytest

from hgswe import cases

import pyoracle


def small_channel():
    return cases.channel(6_000, seed=21, Lx=1500.0, Ly=150.0, Q_in=60.0, nsteps=30)


def wet_dry_case():
    c = cases.channel(6_000, seed=12, Lx=1500.0, Ly=150.0, Q_in=60.0, nsteps=30)
    x = c.mesh.centroid[:, 0]
    h = np.where(x < 500.0, np.maximum(c.h_small, 2.2 - c.zb), c.h_small)
    Q0 = c.Q0.copy()
    Q0[: c.N] = h - c.hstill
    return c.copy_fields(Q0=Q0, dt=cases.cfl_dt(c, 0.4, Q0), name="wet_dry")


def probe_problem(c, nsteps, seed=5, nprobe=25):
    rng = np.random.default_rng(seed)
    probe = rng.choice(c.N, size=nprobe, replace=False).astype(np.int32)
    QT = pyoracle.Oracle(c.copy_fields(n=c.n * 1.15)).run(c.Q0, c.dt, nsteps)
    target = QT[probe] + c.hstill[probe] + c.zb[probe] + rng.normal(0.0, 0.01, nprobe)
    return probe, target


@pytest.mark.parametrize("maker", [small_channel, wet_dry_case])
def test_reverse_gradient_matches_forward_mode(maker):
    c = maker()
    nsteps = 30
    probe, target = probe_problem(c, nsteps)
    orc = pyoracle.Oracle(c)
    loss, gn, gq, g0 = orc.grad_run(c.Q0, c.dt, nsteps, probe, target)
    _, _, _, _, gz = orc.grad_run(c.Q0, c.dt, nsteps, probe, target, wrt=("zb",))
    rng = np.random.default_rng(11)
    for _ in range(2):
        dn = rng.normal(size=c.N) * 1e-3
        dq = rng.normal(size=c.mesh.n_inlet)
        dz = rng.normal(size=c.N) * 1e-2
        _, _, L, dLn = orc.jvp_run(c.Q0, c.dt, nsteps, dn=dn, probe=probe, target=target)
        _, _, _, dLq = orc.jvp_run(c.Q0, c.dt, nsteps, dQin=dq, probe=probe, target=target)
        _, _, _, dLz = orc.jvp_run(c.Q0, c.dt, nsteps, dzb=dz, probe=probe, target=target)
        assert loss == L  # the same forward arithmetic
        assert abs(np.dot(gn, dn) - dLn) <= 1e-10 * abs(dLn)
        assert abs(np.dot(gq, dq) - dLq) <= 1e-10 * abs(dLq)
        assert abs(np.dot(gz, dz) - dLz) <= 1e-10 * abs(dLz)
        dQ0 = rng.normal(size=3 * c.N) * 1e-3
        QN, dQN = orc.jvp_state(c.Q0, dQ0, c.dt, nsteps)
        r = QN[probe] + c.hstill[probe] + c.zb[probe] - target
        dL0 = float(np.sum(2.0 * r * dQN[probe]))
        assert abs(np.dot(g0, dQ0) - dL0) <= 1e-10 * abs(dL0)
    # the gradient is not trivial: the n of every wet cell within reach of a
    # probe over the 30 steps moves the loss (dry cells carry no friction)
    assert np.count_nonzero(gn) > 1000


def test_planted_mistake_is_caught():
    """Seeding the reverse sweep with the gradient of one probe too few
    changes g . v far beyond the tolerance above."""
    c = small_channel()
    nsteps = 30
    probe, target = probe_problem(c, nsteps)
    orc = pyoracle.Oracle(c)
    _, gn, _, _ = orc.grad_run(c.Q0, c.dt, nsteps, probe, target, wrt=("n",))
    _, gn_short, _, _ = orc.grad_run(c.Q0, c.dt, nsteps, probe[1:], target[1:], wrt=("n",))
    dn = np.random.default_rng(3).normal(size=c.N) * 1e-3
    _, _, _, dL = orc.jvp_run(c.Q0, c.dt, nsteps, dn=dn, probe=probe, target=target)
    assert abs(np.dot(gn, dn) - dL) <= 1e-10 * abs(dL)
    assert abs(np.dot(gn_short, dn) - dL) > 1e-4 * abs(dL)


def test_zero_steps_closed_form():
    """With no steps the loss reads the initial state directly: dL/dQ0 is
    2 (xi + hstill + zb - target) at each probe's xi and zero elsewhere,
    dL/dzb the same at the probes, dL/dn zero; a probe outside the mesh is
    refused."""
    c = small_channel()
    probe = np.array([5, 17, 400], np.int32)
    target = np.array([0.5, 1.0, 2.0])
    L, gn, _, g0, gz = pyoracle.Oracle(c).grad_run(c.Q0, c.dt, 0, probe, target, wrt=("n", "Q0", "zb"))
    r = c.Q0[probe] + c.hstill[probe] + c.zb[probe] - target
    assert L == pytest.approx(float(np.sum(r * r)), rel=1e-15)
    want = np.zeros(3 * c.N)
    want[probe] = 2.0 * r
    assert np.array_equal(g0, want)
    assert np.array_equal(gz, want[: c.N])
    assert not np.any(gn)
    with pytest.raises(RuntimeError, match="out of range"):
        pyoracle.Oracle(c).grad_run(c.Q0, c.dt, 1, np.array([c.N], np.int32), np.zeros(1))
