"""GPU parity of the hand-written adjoint (libhgswe's reverse mode).

Pins:
* one RHS evaluation: <ybar, J v> from the oracle's forward-mode (Dual)
  derivative equals <J^T ybar, v> from hg_rhs_vjp, for tangents on Q, the
  per-cell Manning n, the bed zb (bed data rebuilt from zb) and the inlet
  discharge -- exact identity up to rounding (tolerance 1e-10 relative);
* the checkpointed gradient of the probe WSE loss after N steps: g . v
  against the oracle's Dual run of the same N steps (1e-8 relative), every
  component against the oracle's tape reverse mode (1e-9 of the largest),
  and against central finite differences of the GPU forward solve;
* checkpoint-interval invariance and run-to-run determinism (bitwise).
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


def small_channel(seed=21):
    return cases.channel(6_000, seed=seed, Lx=1500.0, Ly=150.0, Q_in=60.0, nsteps=60)


def wet_dry_case():
    c = cases.channel(6_000, seed=12, Lx=1500.0, Ly=150.0, Q_in=60.0, nsteps=60)
    x = c.mesh.centroid[:, 0]
    h = np.where(x < 500.0, np.maximum(c.h_small, 2.2 - c.zb), c.h_small)
    Q0 = c.Q0.copy()
    Q0[: c.N] = h - c.hstill
    return c.copy_fields(Q0=Q0, dt=cases.cfl_dt(c, 0.4, Q0), name="wet_dry")


def corner_inlets_case():
    """Two inlet-Q boundaries (bottom and left walls of a quad channel) that
    share the corner cell, an exit on the right: the inlet adjoint then
    takes the inlets in order in one work-group (a cell on two inlets)."""
    from hgswe import mesh as hm

    xy, cells, sides = cases.rect_quads(200.0, 20.0, 40, 6)
    ns = {1: sides["bottom"], 2: sides["right"], 3: sides["top"], 4: sides["left"]}
    mesh = hm.build_mesh(xy, cells, ns, {1: "INLET-Q", 2: "EXIT-H", 3: "WALL", 4: "INLET-Q"})
    N = mesh.n_cells
    z = np.zeros(N)
    return cases._finish("corner_inlets", mesh, z, np.full(N, 0.03), np.full(N, 1.0), z, z, z, [4.0, 2.0], [1.0],
                         0.02, 50)


def golden(name, tmp_path):
    from conftest import load_golden_case
    from test_oracle_golden import savannah_ic_path

    if name == "bump":
        c = load_golden_case("bump", ic=(0.33, 0.2, 0.0, 0.0))
        return c.copy_fields(dt=0.01)
    c = load_golden_case("savannah", ic=savannah_ic_path(tmp_path))
    return c.copy_fields(dt=0.2 * cases.cfl_dt(c, 1.0))


def vjp_check(c, Q, seed=0, floor=0.0):
    """<ybar, J v> (oracle tangent) against <J^T ybar, v> (device adjoint) for
    tangents on Q, n, zb and the inlet discharges, to 1e-10 of the sum of
    |ybar * J v|; `floor` is an absolute allowance for a derivative that is
    zero up to round-off."""
    ctx, orc = capi.Context(c), pyoracle.Oracle(c)
    rng = np.random.default_rng(seed)
    N = c.N
    ybar = rng.normal(size=3 * N)
    qb, nb, zbb, ib = ctx.rhs_vjp(Q, ybar)
    tangents = {
        "Q": dict(dQ=rng.normal(size=3 * N) * np.concatenate([np.full(N, 1e-2), np.full(2 * N, 1e-2)])),
        "n": dict(dn=rng.normal(size=N) * 1e-3),
        "zb": dict(dzb=rng.normal(size=N) * 1e-2),
    }
    if c.mesh.n_inlet:
        tangents["Qin"] = dict(dQin=rng.normal(size=c.mesh.n_inlet))
    for name, tg in tangents.items():
        _, jv = orc.rhs_jvp(Q, **tg)
        lhs = float(np.dot(ybar, jv))
        v = tg.get("dQ", tg.get("dn", tg.get("dzb", tg.get("dQin"))))
        g = {"Q": qb, "n": nb, "zb": zbb, "Qin": ib}[name]
        rhs = float(np.dot(g, v))
        scale = float(np.sum(np.abs(ybar * jv))) + 1e-300
        say(f"  vjp {c.name} [{name}]: <y,Jv>={lhs:.12e} <J^T y,v>={rhs:.12e} rel={abs(lhs - rhs) / scale:.2e}")
        assert abs(lhs - rhs) <= 1e-10 * scale + floor


def test_rhs_vjp_channel():
    c = small_channel()
    Q = pyoracle.Oracle(c).run(c.Q0, c.dt, 30)
    vjp_check(c, c.Q0, seed=1)
    vjp_check(c, Q, seed=2)


def test_rhs_vjp_wet_dry():
    c = wet_dry_case()
    Q = pyoracle.Oracle(c).run(c.Q0, c.dt, 40)
    h = Q[: c.N] + c.hstill
    say(f"  dry cells: {int(np.sum(h <= c.h_small))}")
    assert np.sum(h <= c.h_small) > 50
    vjp_check(c, Q, seed=3)


def test_rhs_vjp_inlets_sharing_a_cell():
    c = corner_inlets_case()
    gc = c.mesh.ghost_cell[c.mesh.inlet_ghost]
    o = c.mesh.inlet_off
    assert set(gc[o[0]:o[1]].tolist()) & set(gc[o[1]:o[2]].tolist())  # the corner cell
    Q = pyoracle.Oracle(c).run(c.Q0, c.dt, 40)
    vjp_check(c, Q, seed=6)


@pytest.mark.parametrize("name", ["bump", "savannah"])
def test_rhs_vjp_reference_cases(name, tmp_path):
    c = golden(name, tmp_path)
    Q = pyoracle.Oracle(c).run(c.Q0, c.dt, 50)
    vjp_check(c, Q, seed=4)


def _probe_problem(c, nsteps, seed=5, nprobe=30):
    rng = np.random.default_rng(seed)
    probe = rng.choice(c.N, size=nprobe, replace=False).astype(np.int32)
    truth = c.copy_fields(n=c.n * 1.15)
    QT = pyoracle.Oracle(truth).run(c.Q0, c.dt, nsteps)
    target = QT[probe] + c.hstill[probe] + c.zb[probe] + rng.normal(0.0, 0.01, nprobe)  # observation noise
    return probe, target


@pytest.mark.parametrize("maker", [small_channel, wet_dry_case])
def test_gradient_matches_oracle_dual_run(maker):
    c = maker()
    nsteps, kck = 60, 7
    probe, target = _probe_problem(c, nsteps)
    ctx, orc = capi.Context(c), pyoracle.Oracle(c)
    loss, gn, gzb, gq = ctx.grad_probes(c.Q0, c.dt, nsteps, kck, probe, target)
    rng = np.random.default_rng(7)
    dirs = {"n": dict(dn=rng.normal(size=c.N) * 1e-3), "zb": dict(dzb=rng.normal(size=c.N) * 1e-2),
            "Qin": dict(dQin=rng.normal(size=c.mesh.n_inlet))}
    for name, tg in dirs.items():
        _, _, L, dL = orc.jvp_run(c.Q0, c.dt, nsteps, probe=probe, target=target, **tg)
        v = list(tg.values())[0]
        g = {"n": gn, "zb": gzb, "Qin": gq}[name]
        gv = float(np.dot(g, v))
        say(f"  grad {c.name} [{name}]: loss gpu {loss:.12e} oracle {L:.12e}; g.v {gv:.12e} oracle JVP {dL:.12e}")
        assert abs(loss - L) <= 1e-10 * abs(L)
        assert abs(gv - dL) <= 1e-8 * abs(dL)


@pytest.mark.parametrize("maker", [small_channel, wet_dry_case])
def test_gradient_matches_oracle_reverse_mode(maker):
    """Every component of the gradient w.r.t. the per-cell n, the bed zb,
    the inlet discharge and the initial state against the oracle's tape
    reverse mode of the same 30 steps
    (the oracle's gradient is itself pinned against its Dual run in
    test_oracle_reverse.py): max |difference| <= 1e-9 x max |gradient|."""
    c = maker()
    nsteps, kck = 30, 7
    probe, target = _probe_problem(c, nsteps)
    ctx = capi.Context(c)
    loss, gn, gz, gq = ctx.grad_probes(c.Q0, c.dt, nsteps, kck, probe, target)
    orc = pyoracle.Oracle(c)
    L, on, oq, o0 = orc.grad_run(c.Q0, c.dt, nsteps, probe, target)
    _, _, _, _, oz = orc.grad_run(c.Q0, c.dt, nsteps, probe, target, wrt=("zb",))
    # the initial-state cotangent through hg_adjoint, seeded with the loss's
    # final-state cotangent 2 r at the probes
    QN = orc.run(c.Q0, c.dt, nsteps)
    QN_bar = np.zeros(3 * c.N)
    QN_bar[probe] = 2.0 * (QN[probe] + c.hstill[probe] + c.zb[probe] - target)
    g0, gn_adj, _, _ = ctx.adjoint(c.Q0, c.dt, nsteps, kck, QN_bar)
    for name, g, o in (("n", gn, on), ("zb", gz, oz), ("Q0", g0, o0), ("n via hg_adjoint", gn_adj, on)):
        scale = float(np.max(np.abs(o)))
        err = float(np.max(np.abs(g - o)))
        say(f"  grad {c.name} [{name}]: loss gpu {loss:.12e} oracle {L:.12e}; max |diff| {err:.3e} of max {scale:.3e}")
        assert scale > 0.0 and err <= 1e-9 * scale
    say(f"  grad {c.name} [Qin]: {gq} vs {oq}")
    assert abs(loss - L) <= 1e-10 * abs(L)
    assert np.all(np.abs(gq - oq) <= 1e-9 * np.abs(oq))


def test_gradient_matches_oracle_reverse_mode_long_horizon():
    """The same component-wise check over 300 steps of a 10k-cell channel
    with a checkpoint every 25 steps (12 recomputed segments): max
    |difference| <= 1e-9 x max |gradient| for n, and for the initial state
    through hg_adjoint."""
    c = cases.channel(10_000, seed=31, Lx=2000.0, Ly=200.0, Q_in=80.0, nsteps=300)
    nsteps, kck = 300, 25
    probe, target = _probe_problem(c, nsteps, seed=9, nprobe=40)
    ctx = capi.Context(c)
    loss, gn, _, _ = ctx.grad_probes(c.Q0, c.dt, nsteps, kck, probe, target, wrt=("n",))
    orc = pyoracle.Oracle(c)
    L, on, _, o0 = orc.grad_run(c.Q0, c.dt, nsteps, probe, target, wrt=("n", "Q0"))
    pyoracle.Oracle.release_tape()
    QN = orc.run(c.Q0, c.dt, nsteps)
    QN_bar = np.zeros(3 * c.N)
    QN_bar[probe] = 2.0 * (QN[probe] + c.hstill[probe] + c.zb[probe] - target)
    g0, _, _, _ = ctx.adjoint(c.Q0, c.dt, nsteps, kck, QN_bar)
    assert abs(loss - L) <= 1e-10 * abs(L)
    for name, g, o in (("n", gn, on), ("Q0", g0, o0)):
        scale = float(np.max(np.abs(o)))
        err = float(np.max(np.abs(g - o)))
        say(f"  grad {c.name} 300 steps [{name}]: max |diff| {err:.3e} of max {scale:.3e}")
        assert scale > 0.0 and err <= 1e-9 * scale


@pytest.mark.parametrize("maker", [small_channel, wet_dry_case, corner_inlets_case])
def test_tiled_reverse_step_matches_cell_reverse_step(maker):
    """The tiled reverse step (each face's VJP formed once for both cells,
    from the side whose normal points to +x) against the per-cell reverse
    step (each cell forms every face's VJP from its own side): the same
    gradient to rounding.  The tiled gradient is bitwise the same for every
    tile size of the reverse step, whether it shares the step kernel's work
    items or has its own, and every block -> tile mapping (no face's bits
    depend on which work-group evaluates it)."""
    c = maker()
    nsteps, kck = 40, 9
    probe, target = _probe_problem(c, nsteps)
    ref = capi.Context(c, cell_kernel=True).grad_probes(c.Q0, c.dt, nsteps, kck, probe, target)
    ctx = capi.Context(c)
    first = None
    for T in (128, 256, 512):
        for TA in (0, 128, 256, 512):
            for x in ((0, 4) if TA == 0 else (0,)):
                ctx.set_tile_config(T, x)
                ctx.set_adjoint_tile(TA)
                out = ctx.grad_probes(c.Q0, c.dt, nsteps, kck, probe, target)
                if first is None:
                    first = out
                    continue
                assert out[0] == first[0], (T, TA, x)
                for a, b in zip(out[1:], first[1:]):
                    assert np.array_equal(a, b), (T, TA, x)
    assert first[0] == ref[0]  # the forward sweeps are bitwise the same
    for a, b, name in zip(first[1:], ref[1:], ("n", "zb", "Q")):
        rel = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
        say(f"  {c.name} tiled vs per-cell reverse step [{name}]: max rel diff {rel:.2e}")
        assert rel <= 1e-9


def test_unrequested_cotangents_are_skipped_without_changing_the_others():
    """A gradient asked only for d/dn (or only for d/dzb and d/dQ) gives
    bitwise the loss and cotangents of the full request; alternating the
    requests on one context (its cached reversal graphs are kept per
    request) changes nothing."""
    c = small_channel()
    nsteps, kck = 40, 9
    probe, target = _probe_problem(c, nsteps)
    ctx = capi.Context(c)
    full = ctx.grad_probes(c.Q0, c.dt, nsteps, kck, probe, target)
    only_n = ctx.grad_probes(c.Q0, c.dt, nsteps, kck, probe, target, wrt=("n",))
    assert only_n[2] is None and only_n[3] is None
    assert only_n[0] == full[0] and np.array_equal(only_n[1], full[1])
    no_n = ctx.grad_probes(c.Q0, c.dt, nsteps, kck, probe, target, wrt=("zb", "Q"))
    assert no_n[1] is None
    assert np.array_equal(no_n[2], full[2]) and np.array_equal(no_n[3], full[3])
    again = ctx.grad_probes(c.Q0, c.dt, nsteps, kck, probe, target)
    assert again[0] == full[0]
    for a, b in zip(again[1:], full[1:]):
        assert np.array_equal(a, b)


def test_split_adjoint_is_the_adjoint():
    """hg_adjoint_forward then hg_adjoint_reverse (a loss formed from the
    final state between them): the final state is bitwise the device
    forward run's and the cotangents bitwise hg_adjoint's, for a fixed and
    the library's checkpoint interval; the reverse needs its own forward right
    before it (HG_ERR_STATE after another gradient call, a parameter change,
    or a second reverse)."""
    c = small_channel()
    nsteps = 40
    ctx = capi.Context(c)
    lamN = np.random.default_rng(17).normal(size=3 * c.N)
    QN_ref = capi.Context(c).step(c.Q0, c.dt, nsteps)
    for kck in (7, 0):
        ref = ctx.adjoint(c.Q0, c.dt, nsteps, kck, lamN)
        QN = ctx.adjoint_forward(c.Q0, c.dt, nsteps, kck)
        assert np.array_equal(QN, QN_ref)
        out = ctx.adjoint_reverse(lamN)
        for a, b in zip(out, ref):
            assert np.array_equal(a, b), kck
    with pytest.raises(capi.HGError) as e:
        ctx.adjoint_reverse(lamN)
    assert e.value.code == capi.HG_ERR_STATE
    ctx.adjoint_forward(c.Q0, c.dt, nsteps, 7)
    probe = np.arange(10, dtype=np.int32)
    ctx.grad_probes(c.Q0, c.dt, 5, 5, probe, QN_ref[probe])
    with pytest.raises(capi.HGError) as e:
        ctx.adjoint_reverse(lamN)
    assert e.value.code == capi.HG_ERR_STATE
    ctx.adjoint_forward(c.Q0, c.dt, nsteps, 7)
    ctx.set_params(None)
    with pytest.raises(capi.HGError) as e:
        ctx.adjoint_reverse(lamN)
    assert e.value.code == capi.HG_ERR_STATE


def test_checkpoint_interval_invariance_and_determinism():
    c = small_channel()
    nsteps = 40
    probe, target = _probe_problem(c, nsteps)
    ctx = capi.Context(c)
    ref = ctx.grad_probes(c.Q0, c.dt, nsteps, nsteps, probe, target)
    for kck in (1, 7, 13, 0):  # 0: the library's choice (here the whole trajectory)
        out = ctx.grad_probes(c.Q0, c.dt, nsteps, kck, probe, target)
        assert out[0] == ref[0]
        for a, b in zip(out[1:], ref[1:]):
            assert np.array_equal(a, b)


def test_gradient_matches_central_differences():
    """Central differences of the GPU forward solve along random directions
    (the C2-style check at reduced size)."""
    c = cases.channel(20_000, seed=31, Lx=2000.0, Ly=200.0, Q_in=100.0, nsteps=200)
    nsteps, kck = 200, 50
    probe, target = _probe_problem(c, nsteps, nprobe=60)
    ctx = capi.Context(c)
    loss, gn, gzb, gq = ctx.grad_probes(c.Q0, c.dt, nsteps, kck, probe, target)
    rng = np.random.default_rng(11)

    def L(n=None, zb=None, Qin=None):
        cc = c.copy_fields(n=c.n if n is None else n, inletQ=c.inletQ if Qin is None else Qin)
        if zb is not None:
            from hgswe import mesh as hm

            zg, sx, sy = hm.bed_data(c.mesh, zb)
            cc = cc.copy_fields(zb=zb, zb_ghost=zg, S0x=sx, S0y=sy)
        k = capi.Context(cc)
        QN = k.step(c.Q0, c.dt, nsteps)
        r = QN[probe] + cc.hstill[probe] + cc.zb[probe] - target
        return float(np.sum(r * r))

    for trial in range(4):
        vn = rng.normal(size=c.N) * c.n
        e = 1e-5
        fd = (L(n=c.n + e * vn) - L(n=c.n - e * vn)) / (2 * e)
        gv = float(np.dot(gn, vn))
        say(f"  FD n dir {trial}: fd {fd:.10e} adjoint {gv:.10e}")
        assert abs(fd - gv) <= 5e-5 * abs(gv) + 1e-12
    vz = rng.normal(size=c.N) * 0.01
    e = 1e-4
    fd = (L(zb=c.zb + e * vz) - L(zb=c.zb - e * vz)) / (2 * e)
    gv = float(np.dot(gzb, vz))
    say(f"  FD zb: fd {fd:.10e} adjoint {gv:.10e}")
    assert abs(fd - gv) <= 5e-5 * abs(gv) + 1e-12
    e = 1e-3
    fd = (L(Qin=c.inletQ + e) - L(Qin=c.inletQ - e)) / (2 * e)
    say(f"  FD Q: fd {fd:.10e} adjoint {gq[0]:.10e}")
    assert abs(fd - gq[0]) <= 5e-5 * abs(gq[0]) + 1e-12


def test_full_size_adjoint():
    """The 1M-triangle benchmark channel: the RHS adjoint against the oracle's
    tangent (<ybar, J v> = <J^T ybar, v> for tangents on Q, n, zb and the
    inlet discharge), and a 50-step probe-loss gradient that is bitwise the
    same for checkpoint intervals 10 and 25."""
    c = cases.channel(1_000_000, seed=0, nsteps=50)
    ctx = capi.Context(c)
    Q = ctx.step(c.Q0, c.dt, 20)
    vjp_check(c, Q, seed=9)
    probe, target = _probe_problem(c, 50, nprobe=200)
    a = ctx.grad_probes(c.Q0, c.dt, 50, 10, probe, target)
    b = ctx.grad_probes(c.Q0, c.dt, 50, 25, probe, target)
    assert a[0] == b[0]
    for x, y in zip(a[1:], b[1:]):
        assert np.array_equal(x, y)
    say(f"  1M channel gradient: loss {a[0]:.6e}, |dL/dn| {np.linalg.norm(a[1]):.6e}, checkpoint intervals agree")
