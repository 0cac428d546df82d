"""GPU edge cases of the step and its adjoint, against the oracle.

* Cells of degree 5-8, the tiled kernel's 8-slot form. The mesh joins a
  honeycomb (degree 6) and pentagon / heptagon / octagon fans (each polygon
  ringed by triangles) into one block-diagonal mesh, so degrees 3 to 8 meet
  in the same tiles.
* Meshes of one and of two cells.
* Meshes whose last tile holds a single cell: N = T + 1 for each tile size.

Each case checks the RHS against the oracle (1e-12) and steps against the
oracle (1e-10). The tiled kernel must also be bitwise equal to the per-cell
kernel and across tile sizes. The RHS adjoint is checked through the
oracle's tangent, <y, J v> = <J^T y, v> to 1e-10, and probe-loss gradients
against the oracle's Dual run to 1e-8.
"""
import sys

import numpy as np
import pytest

import pyoracle
from hgswe import capi, cases
from hgswe import mesh as hm
from test_gpu_adjoint import vjp_check
from test_gpu_forward import check_rhs, check_steps

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def honeycomb(nx=30, ny=24, R=10.0):
    """Hexagonal cells (pointy-top, offset rows), walls all round, a wavy
    bed and a mound of water."""
    w = np.sqrt(3.0) * R
    key, xy, cells = {}, [], []
    for r in range(ny):
        for c in range(nx):
            cx, cy = w * (c + 0.5 * (r % 2)), 1.5 * R * r
            ids = []
            for j in range(6):
                a = np.pi / 6 + j * np.pi / 3
                p = (cx + R * np.cos(a), cy + R * np.sin(a))
                k = (round(p[0] * 1e6), round(p[1] * 1e6))
                if k not in key:
                    key[k] = len(xy)
                    xy.append(p)
                ids.append(key[k])
            cells.append(ids)
    m = hm.build_mesh(np.array(xy), np.array(cells), {}, {})
    N = m.n_cells
    x, y = m.centroid[:, 0], m.centroid[:, 1]
    zb = 0.1 * np.sin(x / 30.0) * np.cos(y / 40.0)
    xc, yc = x.mean(), y.mean()
    wse = 1.0 + 0.3 * np.exp(-((x - xc) ** 2 + (y - yc) ** 2) / (2 * 40.0 ** 2))
    c = cases._finish("honeycomb", m, zb, np.full(N, 0.03), wse, np.ones(N), np.full(N, 0.05), np.full(N, -0.02),
                      [], [], 0.0, 100)
    return c.copy_fields(dt=cases.cfl_dt(c, 0.4))


def polygon_fan(k, r_in=5.0, r_out=10.0):
    """One k-sided cell ringed by two layers of triangles, walls outside, a
    dam-break start (deep water near the centre)."""
    ang = lambda j, off=0.0: 2 * np.pi * (j + off) / k  # noqa: E731
    inner = [(r_in * np.cos(ang(j)), r_in * np.sin(ang(j))) for j in range(k)]
    outer = [(r_out * np.cos(ang(j, 0.5)), r_out * np.sin(ang(j, 0.5))) for j in range(k)]
    xy = np.array(inner + outer)
    pad = [-1] * (k - 3)
    cells = [list(range(k))]
    for j in range(k):
        cells.append([(j + 1) % k, j, k + j] + pad)
        cells.append([(j + 1) % k, k + j, k + (j + 1) % k] + pad)
    m = hm.build_mesh(xy, np.array(cells), {}, {})
    N = m.n_cells
    r = np.hypot(m.centroid[:, 0], m.centroid[:, 1])
    wse = np.where(r < 3.0, 1.0, 0.4)
    c = cases._finish(f"fan{k}", m, np.zeros(N), np.full(N, 0.03), wse, np.zeros(N), np.zeros(N), np.zeros(N),
                      [], [], 0.0, 100)
    return c.copy_fields(dt=cases.cfl_dt(c, 0.4))


def mixed_polygons():
    return cases.concat([honeycomb(), polygon_fan(5), polygon_fan(7), polygon_fan(8)], name="polygons")


def tiny(ncell):
    """One triangle, or a square split into two triangles; walls; moving water."""
    xy = np.array([[0.0, 0.0], [4.0, 0.0], [4.0, 3.0], [0.0, 3.0]])
    cells = [[0, 1, 2]] if ncell == 1 else [[0, 1, 2], [0, 2, 3]]
    m = hm.build_mesh(xy[: 3 if ncell == 1 else 4], np.array(cells), {}, {})
    N = m.n_cells
    wse = np.array([1.0, 0.7])[:N]
    c = cases._finish(f"tiny{ncell}", m, np.array([0.0, 0.1])[:N], np.full(N, 0.03), wse, np.full(N, 0.5),
                      np.array([0.3, -0.1])[:N], np.array([0.05, 0.2])[:N], [], [], 0.0, 50)
    return c.copy_fields(dt=cases.cfl_dt(c, 0.4))


def strip(ncell, Lx=None):
    """One row of ncell quads, walls, a dam break at mid-length."""
    Lx = Lx or 2.0 * ncell
    xy, cells, _ = cases.rect_quads(Lx, 2.0, ncell, 1)
    m = hm.build_mesh(xy, cells, {}, {})
    N = m.n_cells
    x = m.centroid[:, 0]
    wse = np.where(x < 0.5 * Lx, 1.0, 0.5)
    c = cases._finish(f"strip{ncell}", m, np.zeros(N), np.full(N, 0.03), wse, np.zeros(N), np.zeros(N), np.zeros(N),
                      [], [], 0.0, 60)
    return c.copy_fields(dt=cases.cfl_dt(c, 0.4))


def test_cells_of_degree_five_to_eight():
    c = mixed_polygons()
    deg = np.diff(c.mesh.face_off)
    say(f"  polygons: N={c.N}, degrees {sorted(set(deg.tolist()))}")
    assert deg.max() == 8 and deg.min() == 3 and {5, 6, 7} <= set(deg.tolist())
    ctx, orc = capi.Context(c), pyoracle.Oracle(c)
    check_rhs(c, c.Q0, ctx, orc)
    qg, _ = check_steps(c, c.Q0, c.dt, c.nsteps, ctx, orc)
    check_rhs(c, qg, ctx, orc)
    ref = capi.Context(c, cell_kernel=True).step(c.Q0, c.dt, c.nsteps)
    assert np.array_equal(qg, ref)
    for T in (128, 256, 512):
        ctx.set_tile_config(T, 0)
        assert np.array_equal(ctx.step(c.Q0, c.dt, c.nsteps), ref)


def test_cells_of_degree_five_to_eight_adjoint():
    c = mixed_polygons()
    Q = pyoracle.Oracle(c).run(c.Q0, c.dt, 40)
    vjp_check(c, Q, seed=11)


@pytest.mark.parametrize("ncell", [1, 2])
def test_one_and_two_cell_meshes(ncell):
    c = tiny(ncell)
    assert c.N == ncell
    ctx, orc = capi.Context(c), pyoracle.Oracle(c)
    check_rhs(c, c.Q0, ctx, orc)
    qg, _ = check_steps(c, c.Q0, c.dt, c.nsteps, ctx, orc)
    assert np.array_equal(qg, capi.Context(c, cell_kernel=True).step(c.Q0, c.dt, c.nsteps))
    # one closed cell has no bed-slope derivative (a closed polygon's normals
    # times lengths sum to zero), so its zb products are round-off: both
    # sides must be below 1e-16
    vjp_check(c, qg, seed=12 + ncell, floor=1e-16)
    # probe-loss gradient against the oracle's Dual run
    nsteps, kck = 20, 6
    probe = np.arange(ncell, dtype=np.int32)
    QT = orc.run(c.Q0, c.dt, nsteps)
    # unequal offsets: the two equal-area cells conserve their summed depth,
    # so equal offsets would make the n-derivative of the loss vanish
    target = QT[probe] + c.hstill[probe] + c.zb[probe] + np.array([0.01, -0.02])[:ncell]
    loss, gn, gzb, _ = ctx.grad_probes(c.Q0, c.dt, nsteps, kck, probe, target)
    rng = np.random.default_rng(ncell)
    for name, tg, g in (("n", dict(dn=rng.normal(size=ncell) * 1e-3), gn),
                        ("zb", dict(dzb=rng.normal(size=ncell) * 1e-2), gzb)):
        _, _, L, dL = orc.jvp_run(c.Q0, c.dt, nsteps, probe=probe, target=target, **tg)
        gv = float(np.dot(g, list(tg.values())[0]))
        say(f"  tiny{ncell} grad [{name}]: loss {loss:.12e} oracle {L:.12e}; g.v {gv:.12e} oracle {dL:.12e}")
        assert abs(loss - L) <= 1e-10 * abs(L)
        assert abs(gv - dL) <= 1e-8 * abs(dL)


@pytest.mark.parametrize("T", [128, 256, 512])
def test_last_tile_holds_one_cell(T):
    c = strip(T + 1)
    ctx, orc = capi.Context(c), pyoracle.Oracle(c)
    ctx.set_tile_config(T, 0)
    assert ctx.tile_config()[0] == T
    check_rhs(c, c.Q0, ctx, orc)
    qg, _ = check_steps(c, c.Q0, c.dt, c.nsteps, ctx, orc)
    assert np.array_equal(qg, capi.Context(c, cell_kernel=True).step(c.Q0, c.dt, c.nsteps))


def test_calls_before_fields_are_refused():
    """Every call that launches kernels over the fields returns
    HG_ERR_STATE until hg_set_fields has uploaded them; the context then
    works as usual."""
    c = tiny(2)
    ctx = capi.Context(c, fields=False)
    calls = {
        "advance": lambda: ctx.advance(c.dt, 2),
        "advance_timed": lambda: ctx.advance_timed(c.dt, 2),
        "advance_region": lambda: ctx.advance_region(c.dt, 2),
        "time_advance": lambda: ctx.time_advance(c.dt, 2, 1),
        "rhs": lambda: ctx.rhs(c.Q0),
        "cell_n": lambda: ctx.cell_n(c.Q0),
        "autotune": lambda: ctx.autotune(c.dt, 2),
    }
    for name, call in calls.items():
        with pytest.raises(capi.HGError) as e:
            call()
        say(f"  {name}: {e.value}")
        assert e.value.code == capi.HG_ERR_STATE
    ctx.set_fields(c)
    ref = capi.Context(c).step(c.Q0, c.dt, 5)
    assert np.array_equal(ctx.step(c.Q0, c.dt, 5), ref)


def dry_exit_partial_inlet():
    """A channel whose exit water surface (-1 m) lies below the bed of every
    exit cell, so the exit-H ghost is clamped to h_small (the wet side's
    outflow branches of the Roe flux), and whose inlet reach (x < 100 m) is
    dry over its upper half, so the inlet conveyance and ghost velocities
    carry w = 0 terms.  The deep front against the dry cells makes the
    explicit scheme unstable from the first step on: thin layers with h just
    above h_small take large velocities, and the reference's arithmetic does
    the same.  So the checks take the RHS, one step and its gradient."""
    c = cases.channel(6_000, seed=5, Lx=1500.0, Ly=150.0, Q_in=60.0, nsteps=1)
    N = c.N
    x, y = c.mesh.centroid[:, 0], c.mesh.centroid[:, 1]
    Q0 = c.Q0.copy()
    Q0[:N] = np.where((x < 100.0) & (y > 75.0), c.h_small - c.hstill, Q0[:N])
    c = c.copy_fields(Q0=Q0, exitWSE=np.array([-1.0]), name="dry_exit_partial_inlet")
    return c.copy_fields(dt=cases.cfl_dt(c, 0.4, Q0))


def test_dry_exit_and_partly_dry_inlet():
    c = dry_exit_partial_inlet()
    m = c.mesh
    N = c.N
    exit_cells = m.ghost_cell[m.ghost_kind == hm.BC_EXITH]
    inlet_cells = m.ghost_cell[m.inlet_ghost]
    h0 = c.Q0[:N] + c.hstill
    assert np.all(c.exitWSE[0] - c.zb[exit_cells] < c.h_small)
    assert np.any(h0[inlet_cells] <= c.h_small) and np.any(h0[inlet_cells] > c.h_small)
    ctx, orc = capi.Context(c), pyoracle.Oracle(c)
    check_rhs(c, c.Q0, ctx, orc)
    qg, _ = check_steps(c, c.Q0, c.dt, c.nsteps, ctx, orc)
    assert np.array_equal(qg, capi.Context(c, cell_kernel=True).step(c.Q0, c.dt, c.nsteps))
    vjp_check(c, c.Q0, seed=17)
    # probe-loss gradient against the oracle's Dual run
    nsteps, kck = c.nsteps, 1
    rng = np.random.default_rng(18)
    # after one step n and the inlet discharge reach the water levels only
    # through the inlet cells, so the probes include them (wet and dry)
    probe = np.unique(np.concatenate([rng.choice(N, size=30, replace=False), inlet_cells])).astype(np.int32)
    QT = orc.run(c.Q0, c.dt, nsteps)
    target = QT[probe] + c.hstill[probe] + c.zb[probe] + rng.normal(0.0, 0.01, probe.size)
    loss, gn, gzb, gq = ctx.grad_probes(c.Q0, c.dt, nsteps, kck, probe, target)
    for name, tg, g in (("n", dict(dn=rng.normal(size=N) * 1e-3), gn), ("zb", dict(dzb=rng.normal(size=N) * 1e-2), gzb),
                        ("Qin", dict(dQin=rng.normal(size=m.n_inlet)), gq)):
        _, _, L, dL = orc.jvp_run(c.Q0, c.dt, nsteps, probe=probe, target=target, **tg)
        gv = float(np.dot(g, list(tg.values())[0]))
        say(f"  {c.name} grad [{name}]: g.v {gv:.12e} oracle {dL:.12e}")
        assert abs(loss - L) <= 1e-10 * abs(L)
        assert dL != 0.0
        assert abs(gv - dL) <= 1e-8 * abs(dL)
