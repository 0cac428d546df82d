"""GPU checks of the batched launcher and the launch-shape autotuner.

* A batch built by hg_batch_create from independent instances advances each
  instance exactly as a context of its own does (bitwise), and matches the
  oracle run on the block-diagonal union of the instances.
* Every tiled launch shape (128 / 256 / 512 cells per tile, XCD-contiguous
  or identity mapping) gives bit-identical steps; the autotuner keeps the
  device state it started from.
* The batched gradient of the summed probe loss is the concatenation of the
  per-instance gradients.
"""
import sys

import numpy as np
import pytest

from hgswe import capi, cases

import pyoracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def instances(k=5, seed=40):
    rng = np.random.default_rng(seed)
    out = []
    for r in range(k):
        n = int(rng.integers(1500, 6000))
        out.append(cases.channel(n, seed=seed + r, Lx=1000.0 + 200 * r, Ly=100.0 + 10 * r,
                                 Q_in=float(rng.uniform(20.0, 80.0)), nsteps=60, n_bumps=5))
    out.append(cases.dam_break(40, 6))
    return out


@pytest.fixture(scope="module")
def batch():
    return cases.concat(instances())


def test_batch_matches_instances_and_oracle(batch):
    B = batch
    nsteps = 80
    say(f"  batch: {len(B.extra['instances'])} instances, N={B.N}, inlets={B.mesh.n_inlet}, dt={B.dt}")
    ctx = capi.Context(B)
    QB = ctx.step(B.Q0, B.dt, nsteps)
    QO = pyoracle.Oracle(B).run(B.Q0, B.dt, nsteps)
    rel = float(np.max(np.abs(QB - QO)) / np.max(np.abs(QO)))
    say(f"  batch vs oracle: rel {rel:.2e} bit-identical {float(np.mean(QB == QO)):.4f}")
    assert rel <= 1e-10
    for inst, q in zip(B.extra["instances"], cases.split_state(B, QB)):
        qi = capi.Context(inst).step(inst.Q0, B.dt, nsteps)
        assert np.array_equal(q, qi), inst.name


def test_batch_with_many_inlets_fuses_conveyance():
    """48 inlet boundaries (more than a 32-bit mask holds): per-tile inlet
    lists keep the conveyance fused; results equal the per-cell kernel."""
    inst = [cases.channel(800, seed=100 + r, Lx=600.0, Ly=60.0, Q_in=10.0 + r, nsteps=40, n_bumps=4) for r in range(48)]
    B = cases.concat(inst)
    assert B.mesh.n_inlet == 48
    a = capi.Context(B).step(B.Q0, B.dt, 40)
    b = capi.Context(B, cell_kernel=True).step(B.Q0, B.dt, 40)
    assert np.array_equal(a, b)
    qo = pyoracle.Oracle(B).run(B.Q0, B.dt, 40)
    assert float(np.max(np.abs(a - qo)) / np.max(np.abs(qo))) <= 1e-10


def test_every_tile_shape_is_bit_identical(tmp_path):
    from conftest import load_golden_case
    from test_oracle_golden import savannah_ic_path

    sav = load_golden_case("savannah", ic=savannah_ic_path(tmp_path))
    ch = cases.channel(20_000, seed=11, Lx=2000.0, Ly=200.0, Q_in=100.0, nsteps=40)
    for c, dt in ((ch, ch.dt), (sav, 0.2 * cases.cfl_dt(sav, 1.0))):
        ctx = capi.Context(c)
        ref = ctx.step(c.Q0, dt, 40)
        for T in (128, 256, 512):
            for x in (0, 1, 4, 16, 3000):
                try:
                    ctx.set_tile_config(T, x)
                except capi.HGError as e:  # a tile whose LDS the device does not grant
                    say(f"  {c.name} T={T}: {e}")
                    assert "LDS" in str(e)
                    continue
                assert ctx.tile_config() == (T, x)
                q = ctx.step(c.Q0, dt, 40)
                say(f"  {c.name} T={T} xcd={x}: identical={np.array_equal(q, ref)}")
                assert np.array_equal(q, ref)


def test_autotune_keeps_state(batch):
    B = batch
    ctx = capi.Context(B)
    Q1 = ctx.step(B.Q0, B.dt, 10)
    ctx.upload(Q1)
    T, x, ms = ctx.autotune(B.dt, 32)
    say(f"  autotune: T={T} xcd={x} {ms * 1e3:.2f} us/step")
    assert T in (128, 256, 512) and ms > 0
    assert ctx.tile_config() == (T, x)
    assert np.array_equal(ctx.download(), Q1)
    ctx.advance(B.dt, 20)
    assert np.array_equal(ctx.download(), capi.Context(B).step(Q1, B.dt, 20))


def test_batched_gradient_is_per_instance_gradients(batch):
    B = batch
    inst = B.extra["instances"]
    off = B.extra["cell_off"]
    rng = np.random.default_rng(3)
    nsteps, kck = 40, 9
    probes, targets = [], []
    for r, c in enumerate(inst):
        p = rng.choice(c.N, size=8, replace=False).astype(np.int32)
        probes.append(p)
        targets.append(c.Q0[p] + c.hstill[p] + c.zb[p] + rng.normal(0.0, 0.05, 8))
    probe = np.concatenate([p + off[r] for r, p in enumerate(probes)]).astype(np.int32)
    target = np.concatenate(targets)
    L, gn, gzb, gq = capi.Context(B).grad_probes(B.Q0, B.dt, nsteps, kck, probe, target)
    Ls, iq = 0.0, 0
    for r, c in enumerate(inst):
        l, n_, z_, q_ = capi.Context(c).grad_probes(c.Q0, B.dt, nsteps, kck, probes[r], targets[r])
        Ls += l
        assert np.array_equal(gn[off[r]:off[r + 1]], n_)
        assert np.array_equal(gzb[off[r]:off[r + 1]], z_)
        assert np.array_equal(gq[iq:iq + c.mesh.n_inlet], q_)
        iq += c.mesh.n_inlet
    assert abs(L - Ls) <= 1e-12 * abs(Ls)


def test_per_instance_time_steps(batch):
    """With per-instance steps, each instance advances exactly as it does
    alone with its own dt (forward and gradient)."""
    B = batch
    inst = B.extra["instances"]
    off = B.extra["cell_off"]
    dts = np.array([0.5 * cases.cfl_dt(c, 0.5) for c in inst])
    ctx = capi.Context(B)
    ctx.batch_set_dt(dts)
    nsteps = 30
    qb = ctx.step(B.Q0, 123.0, nsteps)  # the scalar dt is ignored
    for c, q, dt in zip(inst, cases.split_state(B, qb), dts):
        assert np.array_equal(q, capi.Context(c).step(c.Q0, dt, nsteps)), c.name
    rng = np.random.default_rng(8)
    probes, targets = [], []
    for c in inst:
        p = rng.choice(c.N, size=6, replace=False).astype(np.int32)
        probes.append(p)
        targets.append(c.Q0[p] + c.hstill[p] + c.zb[p] + rng.normal(0.0, 0.05, 6))
    probe = np.concatenate([p + off[r] for r, p in enumerate(probes)]).astype(np.int32)
    L, gn, gzb, gq = ctx.grad_probes(B.Q0, 1.0, nsteps, 7, probe, np.concatenate(targets))
    for r, (c, dt) in enumerate(zip(inst, dts)):
        _, n_, z_, _ = capi.Context(c).grad_probes(c.Q0, dt, nsteps, 7, probes[r], targets[r])
        assert np.array_equal(gn[off[r]:off[r + 1]], n_)
        assert np.array_equal(gzb[off[r]:off[r + 1]], z_)
    ctx.batch_set_dt(None)
    assert np.array_equal(ctx.step(B.Q0, B.dt, 5), capi.Context(B).step(B.Q0, B.dt, 5))
