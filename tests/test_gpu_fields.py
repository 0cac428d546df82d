"""K parameter fields on one mesh (hg_fields_create) against the batch of K
copies of the mesh (hg_batch_create), which stores the mesh K times: the
same calls with the same arrays give bitwise the same states, right-hand
sides, gradients and closure weight gradients, for every launch shape and
for the per-cell kernel; the batch itself is pinned to the instances run
alone and to the oracle (tests/test_gpu_batch.py)."""
import sys

import numpy as np
import pytest

from hgswe import capi, cases, ude

import pyoracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _fields(K=4, cells=6_000, seed=3):
    c = cases.channel(cells, seed=seed, Lx=1500.0, Ly=150.0, Q_in=60.0, nsteps=60, n_bumps=6)
    rng = np.random.default_rng(seed)
    inst = []
    for f in range(K):
        n_f = np.exp(rng.normal(np.log(0.035) - 0.045, 0.3, c.N))
        inst.append(c.copy_fields(n=n_f, inletQ=np.array([rng.uniform(20.0, 90.0)]),
                                  exitWSE=c.exitWSE + rng.uniform(-0.2, 0.2)))
    return c, cases.concat(inst)


def test_fields_step_rhs_cfl_bitwise_batch():
    c, B = _fields()
    fb = capi.Context(B, share_mesh=True)
    bb = capi.Context(B)
    say(f"  fields: K=4, N={c.N} per field, tiles {fb.tile_config()}")
    q1 = fb.step(B.Q0, B.dt, 120)
    q2 = bb.step(B.Q0, B.dt, 120)
    assert np.array_equal(q1, q2)
    assert np.array_equal(fb.rhs(q1), bb.rhs(q1))
    assert fb.cfl_dt(q1, 0.5) == bb.cfl_dt(q1, 0.5)
    assert fb.cfl_dt_noalloc(q1, 0.5) == bb.cfl_dt(q1, 0.5)
    # field 2 against the oracle run of that field alone
    inst = B.extra["instances"]
    qo = pyoracle.Oracle(inst[2]).run(inst[2].Q0, B.dt, 120)
    q2f = cases.split_state(B, q1)[2]
    rel = float(np.max(np.abs(q2f - qo)) / np.max(np.abs(qo)))
    say(f"  field 2 after 120 steps vs oracle: rel {rel:.2e}")
    assert rel <= 1e-10


def test_fields_every_launch_shape_and_cell_kernel():
    c, B = _fields(K=3, cells=9_000, seed=5)
    ref = capi.Context(B).step(B.Q0, B.dt, 40)
    fb = capi.Context(B, share_mesh=True)
    for T in (128, 256, 512):
        for x in (0, 1, 4, 16):
            fb.set_tile_config(T, x)
            assert np.array_equal(fb.step(B.Q0, B.dt, 40), ref), (T, x)
    assert np.array_equal(capi.Context(B, share_mesh=True, cell_kernel=True).step(B.Q0, B.dt, 40), ref)


def test_fields_gradient_bitwise_batch():
    c, B = _fields(K=4, cells=5_000, seed=9)
    rng = np.random.default_rng(4)
    off = B.extra["cell_off"]
    probe = np.concatenate([off[f] + rng.choice(c.N, 40, replace=False) for f in range(4)]).astype(np.int32)
    QN = capi.Context(B).step(B.Q0, B.dt, 60)
    target = QN[probe] + B.hstill[probe] + B.zb[probe] + rng.normal(0.0, 0.02, probe.size)
    a = capi.Context(B, share_mesh=True).grad_probes(B.Q0, B.dt, 60, 13, probe, target)
    b = capi.Context(B).grad_probes(B.Q0, B.dt, 60, 13, probe, target)
    assert a[0] == b[0]
    for x, y, name in zip(a[1:], b[1:], ("n", "zb", "Q")):
        say(f"  fields gradient [{name}]: bitwise {bool(np.array_equal(x, y))}, |g| {np.linalg.norm(x):.6e}")
        assert np.array_equal(x, y)
    # the RHS adjoint as well
    ybar = rng.normal(size=3 * B.N)
    ra = capi.Context(B, share_mesh=True).rhs_vjp(QN, ybar)
    rb = capi.Context(B).rhs_vjp(QN, ybar)
    for x, y in zip(ra, rb):
        assert np.array_equal(x, y)


def test_fields_closure_weight_sets_bitwise_batch():
    c, B = _fields(K=4, cells=4_000, seed=12)
    W = ude.perturbed_sets(ude.xavier_weights(2), 4, scale=0.05)
    soc = np.repeat(np.arange(4), c.N).astype(np.int32)
    out = []
    for share in (True, False):
        k = capi.Context(B, share_mesh=share)
        k.set_ude(W, set_of_cell=soc, precision=32)
        q = k.step(B.Q0, B.dt, 50)
        rng = np.random.default_rng(1)
        probe = rng.choice(B.N, 100, replace=False).astype(np.int32)
        target = q[probe] + B.hstill[probe] + B.zb[probe] + 0.01
        g = k.grad_probes(B.Q0, B.dt, 50, 10, probe, target)
        out.append((q, g, k.ude_grad()))
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1][0] == out[1][1][0]
    assert np.array_equal(out[0][2], out[1][2])
    say(f"  closure fields: weight gradient bitwise, |g_w| {np.linalg.norm(out[0][2]):.6e}")


def test_fields_require_shared_mesh_data_and_refuse_what_they_lack():
    c, B = _fields(K=2, cells=2_000, seed=1)
    fb = capi.Context(B, share_mesh=True)
    bad = B.hstill.copy()
    bad[c.N + 5] += 0.1
    with pytest.raises(capi.HGError) as e:
        fb.set_fields(B, hstill=bad)
    assert e.value.code == capi.HG_ERR_ARG
    for kind in (capi.HG_PARAM_ZB, capi.HG_PARAM_N_ZONE):
        with pytest.raises(capi.HGError) as e:
            fb.set_active_param(kind, zone=np.zeros(B.N, np.int32), n_zones=1)
        assert e.value.code == capi.HG_ERR_STATE
    with pytest.raises(capi.HGError) as e:
        fb.batch_set_dt(np.full(2, B.dt))
    assert e.value.code == capi.HG_ERR_STATE
