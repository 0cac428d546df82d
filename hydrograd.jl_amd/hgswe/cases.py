"""Shallow-water cases: the parameter bundle the RHS reads (the fields of
Hydrograd's ``SWE2D_Extra_Parameters``, src/applications/application_commons.jl:7-44)
plus seeded generators for the benchmark workloads and a loader for SRH-2D
cases.

Set-up semantics follow the reference:
  * bed: node z -> cell mean -> zb_ghost / S0 (process_bed_2D.jl:9-66)
  * Manning's n: zone value per cell via the .srhmat zone id
    (process_ManningN_2D.jl:3-98)
  * constant initial condition: hstill = wstill - zb, xi = wse - wstill,
    h = wse - zb (process_ICs_2D.jl:48-61); ghost hstill copies the internal
    cell's (process_ICs_2D.jl:134-169); h < h_small is rejected (:97-99).
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import mesh as hm
from . import srh2d


@dataclass
class Case:
    name: str
    mesh: hm.Mesh2D
    hstill: np.ndarray
    hstill_ghost: np.ndarray
    zb: np.ndarray
    zb_ghost: np.ndarray
    S0x: np.ndarray
    S0y: np.ndarray
    n: np.ndarray
    inletQ: np.ndarray
    exitWSE: np.ndarray
    Q0: np.ndarray
    dt: float
    nsteps: int
    zone: np.ndarray = None
    zone_n: np.ndarray = None
    g: float = 9.81
    k_n: float = 1.0
    h_small: float = 1e-3
    extra: dict = field(default_factory=dict)

    @property
    def N(self):
        return self.mesh.n_cells

    @property
    def G(self):
        return self.mesh.n_ghosts

    def copy_fields(self, **over):
        """A shallow copy with some fields replaced."""
        d = dict(self.__dict__)
        d.update(over)
        return Case(**d)


def _finish(name, mesh, zb_cells, n_cells, wse, wstill, qx, qy, inletQ, exitWSE, dt, nsteps, zone=None,
            zone_n=None, h_small=1e-3, check_ic=True, **extra):
    zb_ghost, S0x, S0y = hm.bed_data(mesh, zb_cells)
    hstill = wstill - zb_cells
    xi = wse - wstill
    h = wse - zb_cells
    if check_ic and np.any(h < h_small):
        raise ValueError("initial water depth below h_small (process_ICs_2D.jl:97-99)")
    Q0 = np.concatenate([xi, qx, qy]).astype(np.float64)
    return Case(name=name, mesh=mesh, hstill=hstill, hstill_ghost=hstill[mesh.ghost_cell].copy(),
                zb=zb_cells, zb_ghost=zb_ghost, S0x=S0x, S0y=S0y, n=n_cells,
                inletQ=np.asarray(inletQ, dtype=np.float64), exitWSE=np.asarray(exitWSE, dtype=np.float64),
                Q0=Q0, dt=float(dt), nsteps=int(nsteps), zone=zone, zone_n=zone_n, h_small=h_small, extra=extra)


# --------------------------------------------------------------------------
def rect_quads(Lx, Ly, nx, ny):
    """Structured quads, CCW, nodes row-major in x."""
    xs = np.linspace(0.0, Lx, nx + 1)
    ys = np.linspace(0.0, Ly, ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing="xy")
    xy = np.stack([X.ravel(), Y.ravel()], axis=1)
    nid = lambda i, j: j * (nx + 1) + i  # noqa: E731
    cells = []
    for j in range(ny):
        for i in range(nx):
            cells.append([nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)])
    cells = np.asarray(cells, dtype=np.int64)
    bottom = [nid(i, 0) for i in range(nx + 1)]
    right = [nid(nx, j) for j in range(ny + 1)]
    top = [nid(i, ny) for i in range(nx, -1, -1)]
    left = [nid(0, j) for j in range(ny, -1, -1)]
    return xy, cells, {"bottom": bottom, "right": right, "top": top, "left": left}


def dam_break(nx=100, ny=10, Lx=200.0, Ly=20.0, n=0.03, dt=0.01, nsteps=200, h_left=1.0, h_right=0.5):
    """Plain dam break on a flat rectangle with walls on all sides: 1,000
    quads, zb = 0, h = 1 m for x < 100 m and 0.5 m otherwise, u = v = 0,
    n = 0.03, dt = 0.01 s, 200 Euler steps."""
    xy, cells, sides = rect_quads(Lx, Ly, nx, ny)
    ns = {1: sides["bottom"], 2: sides["right"], 3: sides["top"], 4: sides["left"]}
    bcs = {1: "WALL", 2: "WALL", 3: "WALL", 4: "WALL"}
    mesh = hm.build_mesh(xy, cells, ns, bcs)
    N = mesh.n_cells
    zb = np.zeros(N)
    wse = np.where(mesh.centroid[:, 0] < Lx / 2, h_left, h_right)
    return _finish("dam_break_1k", mesh, zb, np.full(N, n), wse, np.zeros(N), np.zeros(N), np.zeros(N),
                   [], [], dt, nsteps, zone=np.zeros(N, np.int32), zone_n=np.array([n]),
                   z_nodes=np.zeros(len(xy)), nodestrings=ns, bc_types=bcs)


def jittered_delaunay(nx, ny, Lx, Ly, seed, jitter=0.3):
    """Jittered lattice (interior points moved by up to ``jitter`` spacings,
    boundary points slide along their side), Delaunay-triangulated, CCW."""
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    xs = np.linspace(0.0, Lx, nx)
    ys = np.linspace(0.0, Ly, ny)
    hx, hy = xs[1] - xs[0], ys[1] - ys[0]
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    J = rng.uniform(-jitter, jitter, size=(nx, ny, 2))
    ix = np.arange(nx)[:, None]
    iy = np.arange(ny)[None, :]
    X = X + np.where((ix > 0) & (ix < nx - 1), J[..., 0] * hx, 0.0)
    Y = Y + np.where((iy > 0) & (iy < ny - 1), J[..., 1] * hy, 0.0)
    xy = np.stack([X.ravel(), Y.ravel()], axis=1)
    tri = Delaunay(xy)
    if tri.coplanar.size:
        raise RuntimeError("Delaunay dropped points")
    t = tri.simplices.astype(np.int64)
    p0, p1, p2 = xy[t[:, 0]], xy[t[:, 1]], xy[t[:, 2]]
    cross = (p1[:, 0] - p0[:, 0]) * (p2[:, 1] - p0[:, 1]) - (p1[:, 1] - p0[:, 1]) * (p2[:, 0] - p0[:, 0])
    flip = cross < 0
    t[flip, 1], t[flip, 2] = t[flip, 2].copy(), t[flip, 1].copy()
    keep = np.abs(cross) > 1e-9 * hx * hy
    t = t[keep]
    nid = lambda i, j: i * ny + j  # noqa: E731
    sides = {
        "left": [nid(0, j) for j in range(ny)],            # x = 0, y ascending
        "right": [nid(nx - 1, j) for j in range(ny)],      # x = Lx, y ascending
        "bottom": [nid(i, 0) for i in range(nx)],
        "top": [nid(i, ny - 1) for i in range(nx)],
    }
    return xy, t, sides


def channel(n_target=1_000_000, seed=0, Lx=10_000.0, Ly=1_000.0, Q_in=500.0, wse_out=3.0, n_bumps=50,
            n_zones=5, cfl=0.5, nsteps=1000, zone_n=None, n_cells=None, name=None, h_small=1e-3,
            bc_types=None, wstill_level=None, bump_amp=0.5):
    """Channel case: seeded Delaunay triangulation of a jittered lattice over
    Lx x Ly (~n_target triangles); bed zb = 1e-4 (Lx - x) + 50 Gaussian bumps
    (amplitude U[0, bump_amp] m, bump_amp = 0.5 by default: every cell wet;
    larger amplitudes raise islands above the 3 m surface, with wet/dry
    fronts around them; width U[20, 200] m); Manning n in n_zones
    streamwise zones, each U[0.02, 0.06]; wstill = zb, h = max(h_small,
    3 - zb), q = 0; INLET-Q on x = 0, EXIT-H WSE 3 m on x = Lx, walls
    elsewhere; dt from CFL at the initial state.  ``bc_types`` overrides the
    four sides' types (left, right, bottom, top); ``wstill_level`` sets a
    flat still-water level (hstill = level - zb) instead of wstill = zb."""
    nx = max(3, int(round(math.sqrt(n_target / 2.0 * Lx / Ly))))   # 2236 x 224 lattice for 1M
    ny = max(3, int(round(n_target / 2.0 / nx)))
    xy, tri, sides = jittered_delaunay(nx, ny, Lx, Ly, seed)
    ns = {1: sides["left"], 2: sides["right"], 3: sides["bottom"], 4: sides["top"]}
    bcs = {1: "INLET-Q", 2: "EXIT-H", 3: "WALL", 4: "WALL"}
    if bc_types is not None:
        bcs = {k + 1: t for k, t in enumerate(bc_types)}
    mesh = hm.build_mesh(xy, tri, ns, bcs)
    rng = np.random.default_rng(seed + 1)
    A = rng.uniform(0.0, bump_amp, n_bumps)
    W = rng.uniform(20.0, 200.0, n_bumps)
    cx = rng.uniform(0.0, Lx, n_bumps)
    cy = rng.uniform(0.0, Ly, n_bumps)
    z_nodes = 1e-4 * (Lx - xy[:, 0])
    for k in range(n_bumps):
        r2 = (xy[:, 0] - cx[k]) ** 2 + (xy[:, 1] - cy[k]) ** 2
        z_nodes = z_nodes + A[k] * np.exp(-r2 / (2.0 * W[k] ** 2))
    zb = hm.nodes_to_cells(mesh, z_nodes)
    N = mesh.n_cells
    rng2 = np.random.default_rng(seed + 2)
    zvals = rng2.uniform(0.02, 0.06, n_zones) if zone_n is None else np.asarray(zone_n, dtype=np.float64)
    zone = np.minimum((mesh.centroid[:, 0] / (Lx / n_zones)).astype(np.int32), n_zones - 1) + 1
    n = zvals[zone - 1] if n_cells is None else np.asarray(n_cells, dtype=np.float64)
    h0 = np.maximum(h_small, wse_out - zb)
    wstill = zb.copy() if wstill_level is None else np.full(N, float(wstill_level))
    wse = zb + h0
    for _ in range(4):  # (zb + h_small) - zb may round below h_small; the IC check rejects that
        low = (wse - zb) < h_small
        if not low.any():
            break
        wse = np.where(low, np.nextafter(wse, np.inf), wse)
    n_in = sum(1 for t in bcs.values() if t.upper() == "INLET-Q")
    n_ex = sum(1 for t in bcs.values() if t.upper() == "EXIT-H")
    case = _finish(name or f"channel_{N}", mesh, zb, n, wse, wstill, np.zeros(N), np.zeros(N), [Q_in] * n_in,
                   [wse_out] * n_ex,
                   0.0, nsteps, zone=zone, zone_n=np.concatenate([[zvals[0]], zvals]), h_small=h_small,
                   z_nodes=z_nodes, cfl=cfl, nodestrings=ns, bc_types=bcs)
    case.dt = cfl_dt(case, cfl)
    return case


def cfl_dt(case: Case, cfl: float, Q=None) -> float:
    """dt = CFL * min over wet cells of (2A/P) / (|u| + sqrt(g h)); the same
    operation order as the library's CFL kernel."""
    m = case.mesh
    N = m.n_cells
    Q = case.Q0 if Q is None else Q
    h = Q[:N] + case.hstill
    P = np.zeros(N)
    off = m.face_off.astype(np.int64)
    for k in range(int(m.cell_deg.max())):
        act = k < m.cell_deg
        P = np.where(act, P + m.face_len[np.where(act, off[:-1] + k, 0)], P)
    wet = h > case.h_small
    hh = np.where(wet, h, 1.0)
    u = Q[N:2 * N] / hh
    v = Q[2 * N:] / hh
    spd = np.sqrt(u * u + v * v) + np.sqrt(case.g * hh)
    dti = ((2.0 * m.area) / P) / spd
    return float(cfl * np.min(np.where(wet, dti, np.inf)))


def load_srh_case(case_dir, srhhydro_name, ic=None, dt=0.1, nsteps=10):
    """Load an SRH-2D case the way process_SRH_2D_input.jl:3-225 and
    solve_swe_2D.jl:148-221 do.  ``ic`` is (wse, wstill, qx, qy) constants or
    a path to a JSON with per-cell wse / wstill / q_x / q_y lists."""
    hyd = srh2d.read_srhhydro(os.path.join(case_dir, srhhydro_name))
    geom_name = hyd["Grid"][0]
    mat_name = hyd["HydroMat"][0]
    xyz, cells, ns = srh2d.read_srhgeom(os.path.join(case_dir, geom_name))
    bcs = {k: v[0] for k, v in hyd["BC"].items()}
    mesh = hm.build_mesh(xyz[:, :2], cells, ns, bcs)
    N = mesh.n_cells
    zone, _ = srh2d.read_srhmat(os.path.join(case_dir, mat_name), N)
    nz = max(max(hyd["ManningsN"]) + 1, int(zone.max()) + 1)
    zone_n = np.zeros(nz)
    for k, v in hyd["ManningsN"].items():
        zone_n[k] = float(v[0])
    n = zone_n[zone]
    zb = hm.nodes_to_cells(mesh, xyz[:, 2])
    inlet_ids = mesh.boundary_ids.get(hm.BC_INLETQ, [])
    exit_ids = mesh.boundary_ids.get(hm.BC_EXITH, [])
    inletQ = [float(hyd["IQParams"][b][0]) for b in inlet_ids]
    exitWSE = [float(hyd["EWSParamsC"][b][0]) for b in exit_ids]
    if isinstance(ic, str):
        with open(ic) as fh:
            d = json.load(fh)
        wse = np.asarray(d["wse"], dtype=np.float64)
        wstill = np.asarray(d["wstill"], dtype=np.float64)
        qx = np.asarray(d["q_x"], dtype=np.float64)
        qy = np.asarray(d["q_y"], dtype=np.float64)
    else:
        c = ic if ic is not None else (0.0, 0.0, 0.0, 0.0)
        wse, wstill, qx, qy = (np.full(N, float(v)) for v in c)
    return _finish(os.path.splitext(srhhydro_name)[0], mesh, zb, n, wse, wstill, qx, qy, inletQ, exitWSE, dt,
                   nsteps, zone=zone, zone_n=zone_n, z_nodes=xyz[:, 2])


def concat(instances, name=None, dt=None):
    """The batch of independent cases as one Case over their block-diagonal
    union (instance r's cells, ghosts, inlets, exits follow instance r-1's);
    one dt for all, by default the smallest instance dt."""
    cat = np.concatenate
    m = hm.concat_meshes([c.mesh for c in instances])
    N = [c.N for c in instances]
    Q0 = cat([cat([c.Q0[k * n:(k + 1) * n] for c, n in zip(instances, N)]) for k in range(3)])
    return Case(name=name or f"batch{len(instances)}", mesh=m,
                hstill=cat([c.hstill for c in instances]), hstill_ghost=cat([c.hstill_ghost for c in instances]),
                zb=cat([c.zb for c in instances]), zb_ghost=cat([c.zb_ghost for c in instances]),
                S0x=cat([c.S0x for c in instances]), S0y=cat([c.S0y for c in instances]),
                n=cat([c.n for c in instances]), inletQ=cat([c.inletQ for c in instances]),
                exitWSE=cat([c.exitWSE for c in instances]), Q0=Q0,
                dt=float(min(c.dt for c in instances) if dt is None else dt),
                nsteps=max(c.nsteps for c in instances), g=instances[0].g, k_n=instances[0].k_n,
                h_small=instances[0].h_small, extra={"instances": instances, "cell_off": np.cumsum([0] + N)})


def split_state(batch, Q):
    """Per-instance states [xi; q_x; q_y] from a batch state."""
    off = batch.extra["cell_off"]
    Nt = int(off[-1])
    return [np.concatenate([Q[k * Nt + off[r]: k * Nt + off[r + 1]] for k in range(3)]) for r in range(len(off) - 1)]
