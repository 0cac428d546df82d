"""GPU parity of the forward path: libhgswe (HIP, gfx950) against the CPU
oracle on the same inputs.

Tolerances (the build's contract): the RHS and a single step agree to
1e-12 relative per component (norm-wise, against the component's largest
magnitude); h, hu, hv agree to 1e-10 after N steps.  The kernels repeat the
oracle's operation order without FMA, so most entries are bit-identical;
the only intended difference is the last bit of pow() in the Manning term.
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


def comp_err(a, b, N):
    out = []
    for k in range(3):
        x, y = a[k * N:(k + 1) * N], b[k * N:(k + 1) * N]
        s = np.max(np.abs(y))
        out.append(0.0 if s == 0 else float(np.max(np.abs(x - y)) / s))
    return out


def check_rhs(case, Q, ctx=None, orc=None, tol=1e-12, params=None, recompute_bed=False):
    ctx = ctx or capi.Context(case)
    orc = orc or pyoracle.Oracle(case)
    rg = ctx.rhs(Q, params)
    rc = orc.rhs(Q, recompute_bed=recompute_bed)
    errs = comp_err(rg, rc, case.N)
    same = float(np.mean(rg == rc))
    say(f"  rhs {case.name}: N={case.N} rel err {errs} bit-identical {same:.4f}")
    assert max(errs) <= tol
    return rg, rc


def check_steps(case, Q, dt, nsteps, ctx=None, orc=None, tol=1e-10, params=None, recompute_bed=False):
    ctx = ctx or capi.Context(case)
    orc = orc or pyoracle.Oracle(case)
    qg = ctx.step(Q, dt, nsteps, params)
    qc = orc.run(Q, dt, nsteps, recompute_bed=recompute_bed)
    errs = comp_err(qg, qc, case.N)
    say(f"  {nsteps} steps {case.name}: rel err {errs} bit-identical {float(np.mean(qg == qc)):.4f}")
    assert max(errs) <= tol
    return qg, qc


def test_device_is_gfx950():
    info = capi.info()
    say(info)
    assert "gfx950" in info


def test_dam_break_rhs_and_200_steps():
    c = cases.dam_break()
    ctx, orc = capi.Context(c), pyoracle.Oracle(c)
    check_rhs(c, c.Q0, ctx, orc)
    Q50 = orc.run(c.Q0, c.dt, 50)
    check_rhs(c, Q50, ctx, orc)
    check_steps(c, c.Q0, c.dt, 1, ctx, orc, tol=1e-12)
    check_steps(c, c.Q0, c.dt, c.nsteps, ctx, orc)


@pytest.fixture(scope="module")
def channel20k():
    return cases.channel(20_000, seed=11, Lx=2000.0, Ly=200.0, Q_in=100.0, nsteps=100)


def test_channel_rhs_and_steps(channel20k):
    c = channel20k
    ctx, orc = capi.Context(c), pyoracle.Oracle(c)
    check_rhs(c, c.Q0, ctx, orc)
    Q = orc.run(c.Q0, c.dt, 40)
    check_rhs(c, Q, ctx, orc)
    check_steps(c, c.Q0, c.dt, 1, ctx, orc, tol=1e-12)
    check_steps(c, c.Q0, c.dt, c.nsteps, ctx, orc)


def test_cfl_dt_matches(channel20k):
    c = channel20k
    ctx, orc = capi.Context(c), pyoracle.Oracle(c)
    Q = orc.run(c.Q0, c.dt, 20)
    assert ctx.cfl_dt(Q, 0.5) == orc.cfl_dt(Q, 0.5) == cases.cfl_dt(c, 0.5, Q)


def test_reordering_is_result_neutral(channel20k):
    c = channel20k
    a = capi.Context(c, reorder=True).step(c.Q0, c.dt, 30)
    b = capi.Context(c, reorder=False).step(c.Q0, c.dt, 30)
    assert np.array_equal(a, b)


def test_tiled_kernel_matches_cell_kernel(channel20k, tmp_path):
    """The tiled face-once kernel and the one-work-item-per-cell kernel are
    bit-identical (both reproduce the per-side evaluation exactly)."""
    from conftest import load_golden_case
    from test_oracle_golden import savannah_ic_path

    sav = load_golden_case("savannah", ic=savannah_ic_path(tmp_path))
    for c, dt in ((channel20k, channel20k.dt), (sav, 0.2 * cases.cfl_dt(sav, 1.0))):
        a = capi.Context(c).step(c.Q0, dt, 40)
        b = capi.Context(c, cell_kernel=True).step(c.Q0, dt, 40)
        say(f"  tiled vs cell kernel {c.name}: identical={np.array_equal(a, b)}")
        assert np.array_equal(a, b)
        ra = capi.Context(c).rhs(a)
        rb = capi.Context(c, cell_kernel=True).rhs(a)
        assert np.array_equal(ra, rb)


def test_deterministic_repeat(channel20k):
    c = channel20k
    ctx = capi.Context(c)
    a = ctx.step(c.Q0, c.dt, 25)
    b = ctx.step(c.Q0, c.dt, 25)
    assert np.array_equal(a, b)


def test_wet_dry_fronts():
    """Water released over a bumpy dry bed: exercises the dry-bed Roe
    branches, the h / q clamps and the xi clamp of the step."""
    c = cases.channel(20_000, seed=12, Lx=2000.0, Ly=200.0, bc_types=["WALL"] * 4, nsteps=150)
    x = c.mesh.centroid[:, 0]
    h = np.where(x < 700.0, np.maximum(c.h_small, 2.0 - c.zb), c.h_small)
    Q = c.Q0.copy()
    Q[: c.N] = h - c.hstill
    dt = cases.cfl_dt(c, 0.4, Q)
    ctx, orc = capi.Context(c), pyoracle.Oracle(c)
    check_rhs(c, Q, ctx, orc)
    Qm = orc.run(Q, dt, 60)
    hm = Qm[: c.N] + c.hstill
    say(f"  dry cells after 60 steps: {int(np.sum(hm <= c.h_small))}")
    assert np.sum(hm <= c.h_small) > 100
    check_rhs(c, Qm, ctx, orc)
    check_steps(c, Q, dt, c.nsteps, ctx, orc)


@pytest.mark.parametrize("name", ["bump", "savannah"])
def test_reference_example_cases(name, tmp_path):
    from conftest import load_golden_case
    from test_oracle_golden import savannah_ic_path

    if name == "bump":
        c = load_golden_case("bump", ic=(0.33, 0.2, 0.0, 0.0))
        dt = 0.01
    else:
        c = load_golden_case("savannah", ic=savannah_ic_path(tmp_path))
        dt = 0.2 * cases.cfl_dt(c, 1.0)
    ctx, orc = capi.Context(c), pyoracle.Oracle(c)
    check_rhs(c, c.Q0, ctx, orc)
    Q = orc.run(c.Q0, dt, 100)
    check_rhs(c, Q, ctx, orc)
    check_steps(c, c.Q0, dt, 300, ctx, orc)


def test_active_parameters(channel20k):
    """params_vector per active parameter (semi_discretize_swe_2D.jl:114-199)."""
    c = channel20k
    orc = pyoracle.Oracle(c)
    Q = orc.run(c.Q0, c.dt, 10)
    rng = np.random.default_rng(3)
    # per-cell Manning n
    ctx = capi.Context(c)
    ctx.set_active_param(capi.HG_PARAM_N_CELL)
    n2 = c.n * rng.uniform(0.8, 1.2, c.N)
    orc.set_fields(n=n2)
    check_rhs(c, Q, ctx, orc, params=n2)
    # zone Manning n (zone 0 = default)
    ctx.set_active_param(capi.HG_PARAM_N_ZONE, c.zone, len(c.zone_n))
    zn = c.zone_n * rng.uniform(0.8, 1.2, len(c.zone_n))
    orc.set_fields(n=zn[c.zone])
    check_rhs(c, Q, ctx, orc, params=zn)
    check_steps(c, Q, c.dt, 20, ctx, orc, params=zn)
    # inlet discharge
    ctx.set_active_param(capi.HG_PARAM_Q)
    orc.set_fields(inletQ=np.array([250.0]))
    check_rhs(c, Q, ctx, orc, params=np.array([250.0]))
    # bed elevation: zb_ghost and S0 rebuilt from params (process_bed_2D.jl:46-66)
    ctx.set_active_param(capi.HG_PARAM_ZB)
    zb2 = c.zb + rng.uniform(-0.05, 0.05, c.N)
    orc.set_fields(zb=zb2)
    check_rhs(c, Q, ctx, orc, params=zb2, recompute_bed=True)
    check_steps(c, Q, c.dt, 20, ctx, orc, params=zb2, recompute_bed=True)
    # back to NONE: the fields given to hg_set_fields are used again
    ctx.set_active_param(capi.HG_PARAM_NONE)
    orc2 = pyoracle.Oracle(c)
    check_rhs(c, Q, ctx, orc2)


def test_lake_at_rest_gpu():
    c = cases.channel(200_000, seed=5, bc_types=["WALL"] * 4, wstill_level=3.0)
    r = capi.Context(c).rhs(c.Q0)
    assert np.max(np.abs(r)) < 1e-40 and np.all(r[: c.N] == 0.0)


def test_conveyance_error_is_reported():
    c = cases.channel(5_000, seed=3, Lx=1000.0, Ly=100.0, Q_in=10.0)
    Q = c.Q0.copy()
    inlet_cells = c.mesh.ghost_cell[c.mesh.inlet_ghost]
    Q[inlet_cells] = c.h_small - c.hstill[inlet_cells] - 1.0   # all inlet cells dry
    ctx = capi.Context(c)
    with pytest.raises(capi.HGError) as ei:
        ctx.rhs(Q)
    assert ei.value.code == capi.HG_ERR_CONVEYANCE
    with pytest.raises(RuntimeError):
        pyoracle.Oracle(c).rhs(Q)


def test_full_size_channel_parity():
    """The 1M-triangle channel workload: one RHS and 10 steps against the
    oracle."""
    c = cases.channel(1_000_000, seed=0, nsteps=10)
    say(f"  full-size channel: N={c.N} G={c.G} dt={c.dt}")
    ctx, orc = capi.Context(c), pyoracle.Oracle(c)
    check_rhs(c, c.Q0, ctx, orc)
    check_steps(c, c.Q0, c.dt, 10, ctx, orc)


def test_full_size_thousand_steps_properties():
    """The benchmark workload at full size, 1,000 steps on the device, checked
    through properties that do not need the oracle's run time: the per-cell
    kernel (the reference's cell-by-cell loop shape) and the tiled kernel
    with two launch shapes give bitwise equal states; a repeat is bitwise
    equal; every value is finite; the dry clamp holds (xi >= h_small, and a
    clamped cell carries no discharge)."""
    c = cases.channel(1_000_000, seed=0, nsteps=1000)
    n = 1000
    ctx = capi.Context(c)
    ctx.set_tile_config(256, 16)
    q = ctx.step(c.Q0, c.dt, n)
    assert np.array_equal(ctx.step(c.Q0, c.dt, n), q)
    ctx.set_tile_config(128, 0)
    assert np.array_equal(ctx.step(c.Q0, c.dt, n), q)
    assert np.array_equal(capi.Context(c, cell_kernel=True).step(c.Q0, c.dt, n), q)
    N = c.N
    xi, qx, qy = q[:N], q[N:2 * N], q[2 * N:]
    assert np.all(np.isfinite(q))
    assert np.all(xi >= c.h_small)
    clamped = xi == c.h_small
    assert np.all(qx[clamped] == 0.0) and np.all(qy[clamped] == 0.0)
    moved = float(np.max(np.abs(q - c.Q0)))
    say(f"  1M channel, {n} steps: shapes and per-cell kernel bitwise equal; clamped cells {int(clamped.sum())}; "
        f"max change {moved:.3e}")
    assert moved > 0.0


def test_full_size_wet_dry_fronts():
    """The 1M-triangle channel with islands: bumps up to 4 m under the 3 m
    surface (seed 1, CFL 0.3) put 7.7 % of the cells above the water, with
    wet/dry fronts around them.  One RHS and 100 steps against the oracle;
    1,000 device steps bitwise equal across two launch shapes and the
    per-cell kernel, finite, with the clamp invariant and dry cells left at
    the end; a 50-step gradient bitwise the same for checkpoint intervals 10
    and the library's choice."""
    c = cases.channel(1_000_000, seed=1, nsteps=100, bump_amp=4.0, cfl=0.3)
    N = c.N
    above = float(np.mean(c.zb > 3.0))
    say(f"  islands channel: N={N} dt={c.dt:.5f}, cells above the surface {above:.4f}")
    assert above > 0.05
    ctx, orc = capi.Context(c), pyoracle.Oracle(c)
    check_rhs(c, c.Q0, ctx, orc)
    check_steps(c, c.Q0, c.dt, 100, ctx, orc)
    n = 1000
    ctx.set_tile_config(256, 16)
    q = ctx.step(c.Q0, c.dt, n)
    ctx.set_tile_config(128, 0)
    assert np.array_equal(ctx.step(c.Q0, c.dt, n), q)
    assert np.array_equal(capi.Context(c, cell_kernel=True).step(c.Q0, c.dt, n), q)
    xi, qx, qy = q[:N], q[N:2 * N], q[2 * N:]
    assert np.all(np.isfinite(q))
    assert np.all(xi + c.hstill >= c.h_small)
    clamped = xi + c.hstill == c.h_small
    assert np.all(qx[clamped] == 0.0) and np.all(qy[clamped] == 0.0)
    dry = float(np.mean(clamped))
    say(f"  islands channel, {n} steps: shapes and per-cell kernel bitwise equal; dry fraction {dry:.4f}")
    assert dry > 0.005
    rng = np.random.default_rng(8)
    probe = rng.choice(N, size=200, replace=False).astype(np.int32)
    target = q[probe] + c.hstill[probe] + c.zb[probe] + rng.normal(0.0, 0.01, probe.size)
    a = ctx.grad_probes(c.Q0, c.dt, 50, 10, probe, target)
    b = ctx.grad_probes(c.Q0, c.dt, 50, 0, probe, target)
    assert a[0] == b[0] and all(np.array_equal(x, y) for x, y in zip(a[1:], b[1:]))
    assert all(np.all(np.isfinite(x)) for x in a[1:])
    say(f"  islands channel gradient: loss {a[0]:.6e}, |dL/dn| {np.linalg.norm(a[1]):.6e}")


def decagon_fan():
    """One 10-sided cell ringed by two layers of triangles, walls outside:
    a cell degree the tiled kernel's LDS slots do not take, so the library
    runs the per-cell kernel for it."""
    from hgswe import mesh as hm

    ang = lambda k, off=0.0: 2 * np.pi * (k + off) / 10  # noqa: E731
    inner = [(5 * np.cos(ang(k)), 5 * np.sin(ang(k))) for k in range(10)]
    outer = [(10 * np.cos(ang(k, 0.5)), 10 * np.sin(ang(k, 0.5))) for k in range(10)]
    xy = np.array(inner + outer)
    v = lambda k: k % 10  # noqa: E731
    o = lambda k: 10 + k % 10  # noqa: E731
    cells = [list(range(10))]
    for k in range(10):
        cells.append([v(k + 1), v(k), o(k)] + [-1] * 7)
        cells.append([v(k + 1), o(k), o(k + 1)] + [-1] * 7)
    m = hm.build_mesh(xy, np.array(cells), {}, {})
    N = m.n_cells
    r = np.hypot(m.centroid[:, 0], m.centroid[:, 1])
    wse = np.where(r < 3.0, 1.0, 0.4)
    c = cases._finish("decagon_fan", m, np.zeros(N), np.full(N, 0.03), wse, np.zeros(N), np.zeros(N), np.zeros(N),
                      [], [], 0.0, 200)
    return c.copy_fields(dt=cases.cfl_dt(c, 0.4))


def test_large_polygon_uses_cell_kernel():
    c = decagon_fan()
    assert int(np.max(np.diff(c.mesh.face_off))) == 10
    ctx, orc = capi.Context(c), pyoracle.Oracle(c)
    check_rhs(c, c.Q0, ctx, orc)
    check_steps(c, c.Q0, c.dt, c.nsteps, ctx, orc)
    a = ctx.step(c.Q0, c.dt, 50)
    b = capi.Context(c, cell_kernel=True).step(c.Q0, c.dt, 50)
    assert np.array_equal(a, b)


def test_nonfinite_state_is_reported():
    """A NaN in the state (data, not a fault) ends the call with
    HG_ERR_NONFINITE instead of returning silently corrupted values."""
    c = cases.dam_break(nx=20, ny=4)
    Q = c.Q0.copy()
    Q[7] = np.nan
    ctx = capi.Context(c)
    with pytest.raises(capi.HGError) as e:
        ctx.step(Q, c.dt, 5)
    assert e.value.code == capi.HG_ERR_NONFINITE
    ctx.step(c.Q0, c.dt, 5)  # the context stays usable
