"""The gradient, batch and closure workloads at their full sizes (GPU).

* Adjoint gradient on the 100k-triangle channel, 2,000 steps, checkpoint
  every 50 steps, 200 probes, targets from a lognormal "true" n (mean 0.035,
  sigma 0.3), w.r.t. per-cell n from n = 0.03:
    - against central differences of the device forward solve on 20 random
      directions (Richardson-extrapolated);
    - on a 5k-cell copy of the same generator, against the oracle's
      forward-mode (Dual) derivative along random directions, over the full
      2,000 steps (the reference's Zygote gradient cannot run here; the
      oracle restates its arithmetic).
* 64 uneven instances, 5k-500k triangles each (6.41M cells), one launch per
  step: bitwise equal to three of its instances run alone, 10 steps equal
  to the oracle on the union, the batched gradient w.r.t. n, zb and Q of
  three instances equal to their own gradients, finite with the dry-clamp
  invariant after 400 steps.
* The closure on the 50k-triangle channel, 1,000 steps, 32 f32 weight
  sets batched: two sets bitwise equal to separate runs; 100 f32 closure
  steps against the oracle driven by the NumPy closure; one set's weight
  gradient against central differences.
"""
import sys

import numpy as np
import pytest

from hgswe import capi, cases, ude

import pyoracle
from ude_ref import features, mlp_n

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _lognormal_n(N, rng):
    sig = 0.3
    return np.exp(rng.normal(np.log(0.035) - 0.5 * sig * sig, sig, N))


def _gradient_problem(cells, steps, seed=2):
    c = cases.channel(cells, seed=seed, nsteps=steps)
    rng = np.random.default_rng(seed + 100)
    n_true = _lognormal_n(c.N, rng)
    probe = rng.choice(c.N, size=200, replace=False).astype(np.int32)
    return c, n_true, probe


# ------------------------------------------------------------------ C2 ------
@pytest.fixture(scope="module")
def c2():
    steps = 2000
    c, n_true, probe = _gradient_problem(100_000, steps)
    ctx = capi.Context(c)
    ctx.set_active_param(capi.HG_PARAM_N_CELL)
    QT = ctx.step(c.Q0, c.dt, steps, params=n_true)
    target = QT[probe] + c.hstill[probe] + c.zb[probe]
    say(f"  C2: N={c.N}, dt={c.dt:.6g}, {steps} steps")
    return c, ctx, steps, probe, target


def test_c2_gradient_vs_central_differences_20_directions(c2):
    c, ctx, steps, probe, target = c2
    n0 = np.full(c.N, 0.03)
    loss, gn, _, _ = ctx.grad_probes(c.Q0, c.dt, steps, 50, probe, target, params=n0)

    def L(n):
        QN = ctx.step(c.Q0, c.dt, steps, params=n)
        r = QN[probe] + c.hstill[probe] + c.zb[probe] - target
        return float(np.sum(r * r))

    assert abs(L(n0) - loss) <= 1e-14 * loss  # the gradient's forward sweep is the forward solve
    # Richardson-extrapolated central differences; the difference quotients
    # carry a few 1e-7 of the typical |g.v| of round-off from the 2,000-step
    # solves, so the bar is relative to the RMS of g.v over the directions
    rng = np.random.default_rng(20)
    e = 1e-3
    gvs, fds = [], []
    for d in range(20):
        v = rng.normal(size=c.N) * n0
        D1 = (L(n0 + e * v) - L(n0 - e * v)) / (2 * e)
        D2 = (L(n0 + 0.5 * e * v) - L(n0 - 0.5 * e * v)) / e
        fd = (4.0 * D2 - D1) / 3.0
        gv = float(np.dot(gn, v))
        gvs.append(gv)
        fds.append(fd)
        say(f"  C2 dir {d:2d}: g.v {gv:.12e} FD {fd:.12e} (plain {D2:.12e}) rel {abs(fd - gv) / abs(gv):.1e}")
    gvs, fds = np.array(gvs), np.array(fds)
    scale = float(np.sqrt(np.mean(gvs * gvs)))
    worst = float(np.max(np.abs(fds - gvs)) / scale)
    say(f"  C2: loss {loss:.6e}, |g| {np.linalg.norm(gn):.6e}, max |FD - g.v| / rms(g.v) {worst:.1e}")
    assert worst <= 1e-5


def test_c2_5k_copy_gradient_vs_oracle_dual():
    steps = 2000
    c, n_true, probe = _gradient_problem(5_000, steps)
    orc = pyoracle.Oracle(c.copy_fields(n=n_true))
    pyoracle.set_threads(16)
    QT = orc.run(c.Q0, c.dt, steps)
    target = QT[probe] + c.hstill[probe] + c.zb[probe]
    n0 = np.full(c.N, 0.03)
    ctx = capi.Context(c)
    ctx.set_active_param(capi.HG_PARAM_N_CELL)
    loss, gn, _, _ = ctx.grad_probes(c.Q0, c.dt, steps, 50, probe, target, params=n0)
    o0 = pyoracle.Oracle(c.copy_fields(n=n0))
    rng = np.random.default_rng(5)
    for d in range(3):
        v = rng.normal(size=c.N) * n0
        _, _, Lo, dL = o0.jvp_run(c.Q0, c.dt, steps, dn=v, probe=probe, target=target)
        gv = float(np.dot(gn, v))
        say(f"  C2 5k copy dir {d}: loss gpu {loss:.12e} oracle {Lo:.12e}; g.v {gv:.12e} oracle Dual {dL:.12e}")
        assert abs(loss - Lo) <= 1e-10 * abs(Lo)
        assert abs(gv - dL) <= 1e-8 * abs(dL)
    pyoracle.set_threads(1)


# ------------------------------------------------------------------ C3 ------
def _c3_instances(n_inst=64, seed=7, lo=5_000, hi=500_000):
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), n_inst)).astype(int)
    Qs = rng.uniform(50.0, 1000.0, n_inst)
    return [cases.channel(int(n), seed=1000 + r, Q_in=float(q), nsteps=100) for r, (n, q) in enumerate(zip(sizes, Qs))]


@pytest.fixture(scope="module")
def c3():
    inst = _c3_instances()
    B = cases.concat(inst)
    say(f"  C3: {len(inst)} instances, N={B.N}, inlets={B.mesh.n_inlet}, dt={B.dt:.6g}")
    return inst, B, capi.Context(B)


def _pick3(inst):
    sizes = np.array([c.N for c in inst])
    order = np.argsort(sizes)
    return [int(order[0]), int(order[len(order) // 2]), int(order[-1])]


def test_c3_batch_bitwise_instances_and_oracle(c3):
    inst, B, ctx = c3
    Q10 = ctx.step(B.Q0, B.dt, 10)
    pyoracle.set_threads(16)
    QO = pyoracle.Oracle(B).run(B.Q0, B.dt, 10)
    pyoracle.set_threads(1)
    rel = float(np.max(np.abs(Q10 - QO)) / np.max(np.abs(QO)))
    say(f"  C3 10 steps vs oracle: rel {rel:.2e}, bit-identical fraction {float(np.mean(Q10 == QO)):.5f}")
    assert rel <= 1e-12
    QB = ctx.step(B.Q0, B.dt, 400)
    N = B.N
    assert np.all(np.isfinite(QB))
    xi = QB[:N]
    assert np.all(xi >= B.h_small)  # the dry clamp: xi >= h_small, q = 0 where it clamped
    clamped = xi == B.h_small
    assert np.all(QB[N:2 * N][clamped] == 0.0) and np.all(QB[2 * N:][clamped] == 0.0)
    parts = cases.split_state(B, QB)
    for r in _pick3(inst):
        qi = capi.Context(inst[r]).step(inst[r].Q0, B.dt, 400)
        say(f"  C3 instance {r} (N={inst[r].N}): bitwise {bool(np.array_equal(parts[r], qi))}")
        assert np.array_equal(parts[r], qi)


def test_c3_batched_gradient_is_per_instance_gradients(c3):
    inst, B, ctx = c3
    steps, kck = 200, 50
    off = B.extra["cell_off"]
    rng = np.random.default_rng(11)
    local = [rng.choice(c.N, size=200, replace=False).astype(np.int32) for c in inst]
    probe = np.concatenate([off[r] + p for r, p in enumerate(local)]).astype(np.int32)
    QN = ctx.step(B.Q0, B.dt, steps)
    target = QN[probe] + B.hstill[probe] + B.zb[probe] + rng.normal(0.0, 0.02, probe.size)
    loss, gn, gzb, gq = ctx.grad_probes(B.Q0, B.dt, steps, kck, probe, target)
    assert np.all(np.isfinite(gn)) and np.all(np.isfinite(gzb)) and np.all(np.isfinite(gq))
    for r in _pick3(inst):
        c = inst[r]
        t = target[r * 200:(r + 1) * 200]
        k = capi.Context(c)
        Lr, gnr, gzr, gqr = k.grad_probes(c.Q0, B.dt, steps, kck, local[r], t)
        sl = slice(int(off[r]), int(off[r + 1]))
        say(f"  C3 gradient instance {r} (N={c.N}): n {bool(np.array_equal(gn[sl], gnr))}, "
            f"zb {bool(np.array_equal(gzb[sl], gzr))}, Q {gq[r]!r} vs {gqr[0]!r}")
        assert np.array_equal(gn[sl], gnr)
        assert np.array_equal(gzb[sl], gzr)
        assert gq[r] == gqr[0]


# ------------------------------------------------------------------ C4 ------
@pytest.fixture(scope="module")
def c4():
    c = cases.channel(50_000, seed=4, nsteps=1000)
    W = ude.xavier_weights(0)
    Ws = ude.perturbed_sets(W, 32, scale=0.05)
    return c, W, Ws


def test_c4_32_sets_bitwise_separate_runs(c4):
    c, W, Ws = c4
    nset = 32
    B = cases.concat([c] * nset)
    bctx = capi.Context(B)
    bctx.set_ude(Ws, set_of_cell=np.repeat(np.arange(nset), c.N).astype(np.int32), precision=32)
    QB = bctx.step(B.Q0, B.dt, 1000)
    parts = cases.split_state(B, QB)
    assert np.all(np.isfinite(QB))
    for s in (0, 17):
        k = capi.Context(c)
        k.set_ude(Ws[s], precision=32)
        q = k.step(c.Q0, c.dt, 1000)
        say(f"  C4 set {s}: bitwise {bool(np.array_equal(parts[s], q))}")
        assert np.array_equal(parts[s], q)


def test_c4_f32_closure_steps_vs_oracle(c4):
    """100 steps of the f32 closure against the oracle stepped with the
    NumPy closure's n (f32 network, f64 solver).  The device's f32
    matrix-core sums and NumPy's f32 sums round differently (the network
    output differs by ~1e-7 relative), so the bar is f32-level."""
    c, W, _ = c4
    ctx = capi.Context(c)
    ctx.set_ude(W, precision=32)
    nsteps = 100
    qg = ctx.step(c.Q0, c.dt, nsteps)
    q = c.Q0.copy()
    orc = pyoracle.Oracle(c)
    for _ in range(nsteps):
        orc.set_fields(n=mlp_n(W, features(c, q), dtype=np.float32))
        q = orc.run(q, c.dt, 1)
    N = c.N
    for k, name in enumerate(("xi", "q_x", "q_y")):
        a, b = qg[k * N:(k + 1) * N], q[k * N:(k + 1) * N]
        rel = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
        say(f"  C4 f32 closure {nsteps} steps vs oracle [{name}]: rel {rel:.2e}")
        assert rel <= 1e-5


def test_c4_weight_gradient_fd(c4):
    c, W, _ = c4
    steps, kck = 1000, 50
    rng = np.random.default_rng(8)
    probe = rng.choice(c.N, size=200, replace=False).astype(np.int32)
    ctx = capi.Context(c)
    ctx.set_ude(W, precision=64)
    QN = ctx.step(c.Q0, c.dt, steps)
    target = QN[probe] + c.hstill[probe] + c.zb[probe] + rng.normal(0.0, 0.02, probe.size)
    loss, _, _, _ = ctx.grad_probes(c.Q0, c.dt, steps, kck, probe, target)
    gw = ctx.ude_grad()

    def L(w):
        k = capi.Context(c)
        k.set_ude(w, precision=64)
        Q = k.step(c.Q0, c.dt, steps)
        r = Q[probe] + c.hstill[probe] + c.zb[probe] - target
        return float(np.sum(r * r))

    for d in range(2):
        v = rng.normal(size=W.size)
        e = 1e-4
        fd = (8.0 * (L(W + e * v) - L(W - e * v)) - (L(W + 2 * e * v) - L(W - 2 * e * v))) / (12.0 * e)
        gv = float(np.dot(gw, v))
        say(f"  C4 weights dir {d}: g.v {gv:.10e} FD {fd:.10e}")
        assert abs(fd - gv) <= 1e-5 * abs(gv) + 1e-14
