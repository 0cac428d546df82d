"""GPU checks of the UDE closure n = 0.01 + softplus(MLP(h, |u|, zb)).

* The matrix-core lane layout (16x16x4, f32 and f64) against A @ B.
* The closure's n against a NumPy statement of the same network (f64:
  1e-13 relative; f32: 1e-5).
* The per-cell-n solve is the reference for the step: a step with the closure
  on is bitwise the step of the per-cell-n path given the closure's n, and
  N steps match the oracle driven by the NumPy closure (1e-10).
* Reverse mode through the closure (f64): weight and initial-state
  gradients of the probe loss against central differences; f32 gradients
  against f64 ones.
* Batched instances with one weight set each equal separate runs, forward
  and in the weight gradient of each set; weight sets interleaved within
  blocks: equal sets sum to the one-set gradient, distinct sets against
  central differences; the same two checks on meshes of more than 3,072
  chunks, where a reverse-mode block takes a run of chunks.
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


def small_case(seed=5, n=4000):
    return cases.channel(n, seed=seed, Lx=1200.0, Ly=120.0, Q_in=40.0, nsteps=40, n_bumps=5)


@pytest.mark.parametrize("prec", [32, 64])
def test_mfma_lane_layout(prec):
    rng = np.random.default_rng(prec)
    A, B = rng.normal(size=(16, 4)), rng.normal(size=(4, 16))
    if prec == 32:
        A, B = A.astype(np.float32).astype(np.float64), B.astype(np.float32).astype(np.float64)
    D = capi.mfma_probe(prec, A, B)
    err = float(np.max(np.abs(D - A @ B)))
    say(f"  mfma f{prec}: max err {err:.2e}")
    assert err <= (1e-5 if prec == 32 else 1e-14)


@pytest.mark.parametrize("prec,tol", [(64, 1e-13), (32, 1e-5)])
def test_closure_n_matches_numpy(prec, tol):
    c = small_case()
    Q = pyoracle.Oracle(c).run(c.Q0, c.dt, 30)
    W = ude.xavier_weights(3)
    ctx = capi.Context(c)
    ctx.set_ude(W, precision=prec)
    n = ctx.ude_n(Q)
    ref = mlp_n(W, features(c, Q))
    rel = float(np.max(np.abs(n - ref) / ref))
    say(f"  closure n f{prec}: max rel {rel:.2e}, n in [{n.min():.4f}, {n.max():.4f}]")
    assert rel <= tol


def test_closure_step_is_per_cell_n_step():
    c = small_case()
    Q = pyoracle.Oracle(c).run(c.Q0, c.dt, 20)
    W = ude.xavier_weights(4)
    u = capi.Context(c)
    u.set_ude(W, precision=32)
    n = u.ude_n(Q)
    a = u.step(Q, c.dt, 1)
    p = capi.Context(c.copy_fields(n=n))
    b = p.step(Q, c.dt, 1)
    assert np.array_equal(a, b)
    r1 = u.rhs(Q)
    r2 = p.rhs(Q)
    assert np.array_equal(r1, r2)


def test_closure_steps_match_oracle():
    c = small_case()
    W = ude.xavier_weights(5)
    ctx = capi.Context(c)
    ctx.set_ude(W, precision=64)
    nsteps = 25
    qg = ctx.step(c.Q0, c.dt, nsteps)
    q = c.Q0.copy()
    orc = pyoracle.Oracle(c)
    for _ in range(nsteps):
        orc.set_fields(n=mlp_n(W, features(c, q)))
        q = orc.run(q, c.dt, 1)
    rel = float(np.max(np.abs(qg - q)) / np.max(np.abs(q)))
    say(f"  closure {nsteps} steps vs oracle + NumPy closure: rel {rel:.2e}")
    assert rel <= 1e-10


def test_closure_n_independent_of_chunk_size():
    """A small mesh runs the f32 closure in 32-cell chunks, a large one in
    64-cell chunks; every cell's matrix-core products are the same either
    way, so the n of the same cells is bitwise equal."""
    c = small_case()
    Q = pyoracle.Oracle(c).run(c.Q0, c.dt, 20)
    W = ude.xavier_weights(14)
    one = capi.Context(c)
    one.set_ude(W, precision=32)
    n1 = one.ude_n(Q)
    k = -(-140_000 // c.N)
    B = cases.concat([c] * k)
    assert (B.N + 63) // 64 > 2048 >= (c.N + 63) // 64  # the two chunk sizes
    QB = np.concatenate([Q.reshape(3, c.N)] * k, axis=1).reshape(-1)
    big = capi.Context(B)
    big.set_ude(W, precision=32)
    nb = big.ude_n(QB)
    for r in (0, k // 2, k - 1):
        assert np.array_equal(nb[r * c.N:(r + 1) * c.N], n1), r


def fd4(f, x, v, e):
    """Fourth-order central difference of f at x along v."""
    return (8.0 * (f(x + e * v) - f(x - e * v)) - (f(x + 2 * e * v) - f(x - 2 * e * v))) / (12.0 * e)


def _probe(c, nsteps, seed=9, k=25):
    rng = np.random.default_rng(seed)
    probe = rng.choice(c.N, size=k, replace=False).astype(np.int32)
    Q = pyoracle.Oracle(c).run(c.Q0, c.dt, nsteps)
    target = Q[probe] + c.hstill[probe] + c.zb[probe] + rng.normal(0.0, 0.02, k)
    return probe, target


def test_closure_weight_gradient_fd():
    c = small_case()
    nsteps, kck = 30, 7
    probe, target = _probe(c, nsteps)
    W = ude.xavier_weights(6)
    ctx = capi.Context(c)
    ctx.set_ude(W, precision=64)
    loss, _, _, _ = ctx.grad_probes(c.Q0, c.dt, nsteps, kck, probe, target)
    gw = ctx.ude_grad()
    again = capi.Context(c)
    again.set_ude(W, precision=64)
    again.grad_probes(c.Q0, c.dt, nsteps, kck, probe, target)
    assert np.array_equal(gw, again.ude_grad())  # deterministic

    def L(w):
        k = capi.Context(c)
        k.set_ude(w, precision=64)
        q = k.step(c.Q0, c.dt, nsteps)
        r = q[probe] + c.hstill[probe] + c.zb[probe] - target
        return float(np.sum(r * r))

    rng = np.random.default_rng(2)
    for trial in range(3):
        # the loss moves little with the weights, so a small step drowns in
        # round-off; a fourth-order stencil with a large step is accurate
        v = rng.normal(size=W.size)
        fd = fd4(L, W, v, 5e-3)
        gv = float(np.dot(gw, v))
        say(f"  closure weights dir {trial}: fd {fd:.10e} adjoint {gv:.10e} rel {abs(fd - gv) / abs(gv):.2e}")
        assert abs(fd - gv) <= 1e-5 * abs(gv)  # difference-quotient round-off and truncation at this step
    # initial-state gradient through the closure
    lb = np.zeros(3 * c.N)
    rb = 2.0 * (ctx.step(c.Q0, c.dt, nsteps)[probe] + c.hstill[probe] + c.zb[probe] - target)
    np.add.at(lb, probe, rb)
    Q0b, _, _, _ = ctx.adjoint(c.Q0, c.dt, nsteps, kck, lb)
    v = np.zeros(3 * c.N)
    v[: c.N] = rng.normal(size=c.N) * 1e-2

    def Lq(q0):
        k = capi.Context(c)
        k.set_ude(W, precision=64)
        q = k.step(q0, c.dt, nsteps)
        r = q[probe] + c.hstill[probe] + c.zb[probe] - target
        return float(np.sum(r * r))

    fd = fd4(Lq, c.Q0, v, 0.05)
    gv = float(np.dot(Q0b, v))
    say(f"  closure Q0 dir: fd {fd:.10e} adjoint {gv:.10e} rel {abs(fd - gv) / abs(gv):.2e}")
    assert abs(fd - gv) <= 1e-5 * abs(gv)


def test_closure_gradient_independent_of_checkpoint_interval():
    """The recompute keeps each step's closure n for the reverse step.  The
    reverse steps see the same states and n whatever the segmenting, and each
    block adds its weight cotangents in step order, so the gradient is bitwise
    the same for any checkpoint interval (a wrongly indexed n would differ)."""
    c = small_case()
    nsteps = 23
    probe, target = _probe(c, nsteps)
    W = ude.xavier_weights(13)
    res = []
    for kck in (1, 4, 23):
        ctx = capi.Context(c)
        ctx.set_ude(W, precision=32)
        loss, _, _, _ = ctx.grad_probes(c.Q0, c.dt, nsteps, kck, probe, target)
        Q0b, _, _, _ = ctx.adjoint(c.Q0, c.dt, nsteps, kck, np.ones(3 * c.N))
        res.append((loss, ctx.ude_grad(), Q0b))
    for kck, r in zip((4, 23), res[1:]):
        assert r[0] == res[0][0], kck
        assert np.array_equal(r[1], res[0][1]), kck
        assert np.array_equal(r[2], res[0][2]), kck
    assert np.linalg.norm(res[0][1]) > 0.0


def test_closure_f32_gradient_tracks_f64():
    c = small_case()
    nsteps, kck = 30, 10
    probe, target = _probe(c, nsteps)
    W = ude.xavier_weights(7)
    g = {}
    for prec in (32, 64):
        ctx = capi.Context(c)
        ctx.set_ude(W, precision=prec)
        ctx.grad_probes(c.Q0, c.dt, nsteps, kck, probe, target)
        g[prec] = ctx.ude_grad()
    rel = float(np.linalg.norm(g[32] - g[64]) / np.linalg.norm(g[64]))
    say(f"  closure gradient f32 vs f64: rel {rel:.2e}")
    assert rel <= 1e-3


def test_batched_weight_sets_equal_separate_runs():
    inst = [small_case(seed=s, n=2500) for s in (1, 2, 3)]
    B = cases.concat(inst)
    Ws = ude.perturbed_sets(ude.xavier_weights(8), 3, scale=0.1)
    set_of_cell = np.repeat(np.arange(3), [c.N for c in inst]).astype(np.int32)
    ctx = capi.Context(B)
    ctx.set_ude(Ws, set_of_cell=set_of_cell, precision=32)
    qb = ctx.step(B.Q0, B.dt, 30)
    for r, (c, q) in enumerate(zip(inst, cases.split_state(B, qb))):
        k = capi.Context(c)
        k.set_ude(Ws[r], precision=32)
        assert np.array_equal(q, k.step(c.Q0, B.dt, 30)), r


def test_batched_weight_set_gradients_equal_separate_runs():
    """One reverse sweep over a batch with one weight set per instance gives
    each set the gradient of its own instance's probe loss.  The per-block
    partial sums group the cells differently from a run alone, so the match
    is to round-off, not bitwise."""
    inst = [small_case(seed=s, n=n) for s, n in ((1, 2500), (2, 1900), (3, 3100))]
    B = cases.concat(inst)
    Ws = ude.perturbed_sets(ude.xavier_weights(8), 3, scale=0.1)
    set_of_cell = np.repeat(np.arange(3), [c.N for c in inst]).astype(np.int32)
    nsteps, kck = 20, 6
    probes, targets = [], []
    for r, c in enumerate(inst):
        p, t = _probe(c, 0, seed=20 + r, k=15)
        probes.append(p)
        targets.append(t)
    ctx = capi.Context(B)
    ctx.set_ude(Ws, set_of_cell=set_of_cell, precision=64)
    first = np.cumsum([0] + [c.N for c in inst[:-1]])
    bprobe = np.concatenate([p + o for p, o in zip(probes, first)]).astype(np.int32)
    ctx.grad_probes(B.Q0, B.dt, nsteps, kck, bprobe, np.concatenate(targets))
    gb = ctx.ude_grad()
    assert gb.shape == (3, capi.UDE_WEIGHTS)
    for r, c in enumerate(inst):
        k = capi.Context(c)
        k.set_ude(Ws[r], precision=64)
        k.grad_probes(c.Q0, B.dt, nsteps, kck, probes[r], targets[r])
        g1 = k.ude_grad()
        rel = float(np.max(np.abs(gb[r] - g1)) / np.max(np.abs(g1)))
        say(f"  set {r}: |g| {np.linalg.norm(g1):.4e}, batch vs alone max rel {rel:.2e}")
        assert np.linalg.norm(g1) > 0.0
        assert rel <= 1e-12, r


def _interleaved(c, nsets):
    """Each cell's set is its caller index mod nsets, so every 64-cell block
    holds several sets whatever the cell order."""
    return (np.arange(c.N) % nsets).astype(np.int32)


def test_interleaved_weight_sets_sum_to_single_set_gradient():
    c = small_case(n=3000)
    nsteps, kck = 20, 5
    probe, target = _probe(c, nsteps)
    W = ude.xavier_weights(11)
    one = capi.Context(c)
    one.set_ude(W, precision=64)
    one.grad_probes(c.Q0, c.dt, nsteps, kck, probe, target)
    g1 = one.ude_grad()
    many = capi.Context(c)
    many.set_ude(np.stack([W] * 4), set_of_cell=_interleaved(c, 4), precision=64)
    qa = many.step(c.Q0, c.dt, nsteps)
    assert np.array_equal(qa, one.step(c.Q0, c.dt, nsteps))  # equal weights: the same trajectory
    many.grad_probes(c.Q0, c.dt, nsteps, kck, probe, target)
    g4 = many.ude_grad()
    assert g4.shape[0] == 4 and all(np.linalg.norm(g) > 0.0 for g in g4)
    rel = float(np.max(np.abs(g4.sum(axis=0) - g1)) / np.max(np.abs(g1)))
    say(f"  4 interleaved equal sets: sum vs one set max rel {rel:.2e}")
    assert rel <= 1e-12


def test_interleaved_weight_set_gradient_fd():
    """Distinct interleaved sets: the gradient of one set's weights against
    central differences that move only that set."""
    c = small_case(n=3000)
    nsteps, kck = 20, 5
    probe, target = _probe(c, nsteps)
    Ws = ude.perturbed_sets(ude.xavier_weights(12), 3, scale=0.2)
    soc = _interleaved(c, 3)
    ctx = capi.Context(c)
    ctx.set_ude(Ws, set_of_cell=soc, precision=64)
    ctx.grad_probes(c.Q0, c.dt, nsteps, kck, probe, target)
    g = ctx.ude_grad()

    def L(ws):
        k = capi.Context(c)
        k.set_ude(ws, set_of_cell=soc, precision=64)
        q = k.step(c.Q0, c.dt, nsteps)
        r = q[probe] + c.hstill[probe] + c.zb[probe] - target
        return float(np.sum(r * r))

    rng = np.random.default_rng(3)
    for s in (0, 2):
        v = np.zeros_like(Ws)
        v[s] = rng.normal(size=Ws.shape[1])
        fd = fd4(L, Ws, v, 5e-3)
        gv = float(np.dot(g[s], v[s]))
        say(f"  interleaved set {s}: fd {fd:.10e} adjoint {gv:.10e} rel {abs(fd - gv) / abs(gv):.2e}")
        assert abs(fd - gv) <= 1e-5 * abs(gv)


def test_unused_weight_set_has_zero_cotangent():
    """A set that no cell uses has an empty run of slots: its cotangent is
    exactly zero, the used sets' are not, and a set's weights do not act
    where it is unused."""
    c = small_case(n=3000)
    nsteps, kck = 10, 4
    probe, target = _probe(c, nsteps)
    Ws = ude.perturbed_sets(ude.xavier_weights(15), 4, scale=0.2)
    soc = np.where(np.arange(c.N) % 2 == 0, 0, 3).astype(np.int32)  # sets 1 and 2 unused
    ctx = capi.Context(c)
    ctx.set_ude(Ws, set_of_cell=soc, precision=64)
    ctx.grad_probes(c.Q0, c.dt, nsteps, kck, probe, target)
    g = ctx.ude_grad()
    assert g.shape == (4, capi.UDE_WEIGHTS)
    assert not np.any(g[1]) and not np.any(g[2])
    assert np.linalg.norm(g[0]) > 0.0 and np.linalg.norm(g[3]) > 0.0
    W2 = Ws.copy()
    W2[1:3] = ude.xavier_weights(16)  # change only the unused sets
    k = capi.Context(c)
    k.set_ude(W2, set_of_cell=soc, precision=64)
    assert np.array_equal(k.step(c.Q0, c.dt, nsteps), ctx.step(c.Q0, c.dt, nsteps))


def test_chunk_run_weight_gradients_equal_separate_runs():
    """Above 3,072 64-cell chunks a reverse-mode block of the closure takes a
    run of consecutive chunks and keeps adding each set's weight cotangents
    across them.  Seven instances of 29,336 cells (3,208 chunks: two per
    block, some blocks across two instances) give each set the gradient of its own
    instance's loss, which a run alone (one chunk per block) groups
    differently: equal to round-off in f64, and the f32 network tracks it."""
    inst = [small_case(seed=s, n=30_000) for s in range(1, 8)]
    B = cases.concat(inst)
    assert (B.N + 63) // 64 > 3072
    Ws = ude.perturbed_sets(ude.xavier_weights(21), len(inst), scale=0.1)
    set_of_cell = np.repeat(np.arange(len(inst)), [c.N for c in inst]).astype(np.int32)
    nsteps, kck = 12, 4
    probes, targets = [], []
    for r, c in enumerate(inst):
        p, t = _probe(c, 0, seed=40 + r, k=15)
        probes.append(p)
        targets.append(t)
    first = np.cumsum([0] + [c.N for c in inst[:-1]])
    bprobe = np.concatenate([p + o for p, o in zip(probes, first)]).astype(np.int32)
    g = {}
    for prec in (64, 32):
        ctx = capi.Context(B)
        ctx.set_ude(Ws, set_of_cell=set_of_cell, precision=prec)
        ctx.grad_probes(B.Q0, B.dt, nsteps, kck, bprobe, np.concatenate(targets))
        g[prec] = ctx.ude_grad()
    for r, c in enumerate(inst):
        k = capi.Context(c)
        k.set_ude(Ws[r], precision=64)
        k.grad_probes(c.Q0, B.dt, nsteps, kck, probes[r], targets[r])
        g1 = k.ude_grad()
        rel = float(np.max(np.abs(g[64][r] - g1)) / np.max(np.abs(g1)))
        rel32 = float(np.linalg.norm(g[32][r] - g1) / np.linalg.norm(g1))
        say(f"  set {r}: runs of chunks vs alone max rel {rel:.2e}; f32 vs f64 {rel32:.2e}")
        assert np.linalg.norm(g1) > 0.0
        assert rel <= 1e-12, r
        assert rel32 <= 1e-3, r


def test_chunk_run_interleaved_sets_sum_to_single_set_gradient():
    """Runs of chunks with four equal sets interleaved cell by cell: every
    pass switches set, so each block adds into each of its four slots many
    times; the four gradients sum to the one-set gradient of the same mesh."""
    c = small_case(seed=6, n=210_000)
    assert (c.N + 63) // 64 > 3072
    nsteps, kck = 8, 4
    probe, target = _probe(c, 0, seed=11, k=30)
    W = ude.xavier_weights(22)
    one = capi.Context(c)
    one.set_ude(W, precision=64)
    one.grad_probes(c.Q0, c.dt, nsteps, kck, probe, target)
    g1 = one.ude_grad()
    many = capi.Context(c)
    many.set_ude(np.stack([W] * 4), set_of_cell=_interleaved(c, 4), precision=64)
    many.grad_probes(c.Q0, c.dt, nsteps, kck, probe, target)
    g4 = many.ude_grad()
    assert all(np.linalg.norm(g) > 0.0 for g in g4)
    rel = float(np.max(np.abs(g4.sum(axis=0) - g1)) / np.max(np.abs(g1)))
    say(f"  runs of chunks, 4 interleaved equal sets: sum vs one set max rel {rel:.2e}")
    assert rel <= 1e-12


def test_closure_rejects_n_as_active_parameter():
    c = small_case()
    ctx = capi.Context(c)
    ctx.set_ude(ude.xavier_weights(1), precision=32)
    with pytest.raises(capi.HGError):
        ctx.set_active_param(capi.HG_PARAM_N_CELL)
    ctx.set_ude(None)
    ctx.set_active_param(capi.HG_PARAM_N_CELL)
