"""Flattened unstructured mesh: the data of Hydrograd's ``mesh_2D`` and
``BoundaryConditions2D`` that the shallow-water RHS reads, in the
cell-ordered CSR form the HIP library consumes.

Reference structures (Hydrograd.jl):
  * ``mesh_2D`` struct and ``initialize_mesh_2D`` -- src/meshes/mesh_2D.jl:2-71, 75-453
  * edges / boundary edges                     -- src/utilities/SRH_2D/SRH_2D_SRHGeom.jl:251-382
  * geometry (shoelace area, centroid, normals) -- src/meshes/mesh_2D.jl:456-652
  * boundary lists per BC type                 -- src/fvm/boundary_conditions/bc_2D.jl:154-570

Conventions (0-based, cells in the caller's order):
  * cell i owns CSR slots ``face_off[i] .. face_off[i+1]-1``, one per local
    face k = edge node_k -> node_{k+1}, in the reference's local order;
  * ``face_nbr[s]`` is the neighbouring cell, or ``~g`` (negative) for ghost g;
  * ``face_nx/face_ny`` are ``cell_normals[i][k]`` and ``face_len`` is
    ``face_lengths[edge]``;
  * every boundary edge has exactly one ghost; ghosts are numbered in
    increasing edge order (the reference numbers them in Dict iteration
    order, which no per-cell output depends on).

Every float is computed with the reference's operation order, so the arrays
are bit-identical to what ``initialize_mesh_2D`` produces for the same nodes.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

BC_INLETQ, BC_EXITH, BC_WALL, BC_SYMM = 0, 1, 2, 3
_BC_NAMES = {"inlet-q": BC_INLETQ, "exit-h": BC_EXITH, "wall": BC_WALL, "symm": BC_SYMM}


def bc_kind(name: str) -> int:
    """Boundary type string of a .srhhydro ``BC`` line -> kind code
    (bc_2D.jl:176-239 compares the lower-cased names)."""
    key = name.strip().lower()
    if key not in _BC_NAMES:
        raise ValueError(f"unsupported boundary type {name!r}")
    return _BC_NAMES[key]


@dataclass
class Mesh2D:
    node_xy: np.ndarray          # (nNodes, 2)
    cell_nodes: np.ndarray       # (N, maxdeg) 0-based, padded with -1
    cell_deg: np.ndarray         # (N,)
    edge_nodes: np.ndarray       # (F, 2) sorted node pair of each edge
    face_off: np.ndarray         # (N+1,) int32
    face_nbr: np.ndarray         # (M,) int32
    face_edge: np.ndarray        # (M,) int32
    face_nx: np.ndarray          # (M,)
    face_ny: np.ndarray
    face_len: np.ndarray
    area: np.ndarray             # (N,)
    centroid: np.ndarray         # (N, 2)
    # boundaries
    ghost_edge: np.ndarray       # (G,)
    ghost_cell: np.ndarray       # (G,) int32
    ghost_slot: np.ndarray       # (G,) int32 CSR slot of the boundary face
    ghost_kind: np.ndarray       # (G,) int32
    ghost_bc: np.ndarray         # (G,) int32 index among inlets / exits, else -1
    ghost_nx: np.ndarray         # (G,) face_direction .* face_normals[edge]
    ghost_ny: np.ndarray
    ghost_len: np.ndarray        # (G,) face_lengths[edge]
    n_inlet: int
    n_exit: int
    inlet_off: np.ndarray        # (n_inlet+1,) int32
    inlet_ghost: np.ndarray      # ghosts of each inlet, NodeString order
    boundary_ids: dict = field(default_factory=dict)  # kind -> list of boundary ids

    @property
    def n_cells(self) -> int:
        return int(self.cell_deg.shape[0])

    @property
    def n_ghosts(self) -> int:
        return int(self.ghost_cell.shape[0])

    @property
    def n_faces(self) -> int:
        return int(self.edge_nodes.shape[0])


def _polygon_geometry(xy: np.ndarray, cell_nodes: np.ndarray, deg: np.ndarray):
    """Shoelace area and centroid, sequential over vertices from 0.0
    (mesh_2D.jl:600-634)."""
    N, maxdeg = cell_nodes.shape
    area = np.zeros(N)
    cx = np.zeros(N)
    cy = np.zeros(N)
    rows = np.arange(N)
    for k in range(maxdeg):
        active = k < deg
        i1 = cell_nodes[:, k]
        i2 = cell_nodes[rows, np.where(k + 1 < deg, k + 1, 0)]
        x1, y1 = xy[np.where(active, i1, 0), 0], xy[np.where(active, i1, 0), 1]
        x2, y2 = xy[np.where(active, i2, 0), 0], xy[np.where(active, i2, 0), 1]
        cross = x1 * y2 - x2 * y1
        area = np.where(active, area + cross, area)
        cx = np.where(active, cx + (x1 + x2) * cross, cx)
        cy = np.where(active, cy + (y1 + y2) * cross, cy)
    if np.any(area < 0):
        bad = int(np.nonzero(area < 0)[0][0])
        raise ValueError(f"cell {bad + 1} is clockwise (mesh_2D.jl:591-613 rejects it)")
    a = np.abs(area) / 2
    cen = np.stack([cx / (6 * a), cy / (6 * a)], axis=1)
    return a, cen


def build_mesh(node_xy, cells, nodestrings: dict, bc_types: dict) -> Mesh2D:
    """Build the flattened mesh.

    node_xy     : (nNodes, 2) coordinates
    cells       : (N, maxdeg) 0-based node ids padded with -1, CCW
    nodestrings : {boundary id (1-based): sequence of 0-based node ids}
    bc_types    : {boundary id: "INLET-Q" | "EXIT-H" | "WALL" | "SYMM"}

    Boundary edges not covered by a NodeString with a BC go to a default wall
    boundary with id ``len(assigned) + 1`` (SRH_2D_SRHGeom.jl:345-349,
    mesh_2D.jl:173-181).
    """
    xy = np.ascontiguousarray(node_xy, dtype=np.float64)[:, :2]
    cell_nodes = np.asarray(cells, dtype=np.int64)
    deg = (cell_nodes >= 0).sum(axis=1).astype(np.int32)
    N, maxdeg = cell_nodes.shape
    nn = xy.shape[0]

    # ---- slots (cell, local face) and edges in first-appearance order -------
    face_off = np.zeros(N + 1, dtype=np.int64)
    face_off[1:] = np.cumsum(deg)
    M = int(face_off[-1])
    slot_cell = np.repeat(np.arange(N), deg)
    slot_k = np.arange(M) - face_off[slot_cell]
    a = cell_nodes[slot_cell, slot_k]
    b = cell_nodes[slot_cell, np.where(slot_k + 1 < deg[slot_cell], slot_k + 1, 0)]
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    key = lo * nn + hi
    uniq, first, inv = np.unique(key, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")          # edge ids by first appearance
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    face_edge = rank[inv].astype(np.int64)
    F = uniq.size
    edge_nodes = np.stack([uniq[order] // nn, uniq[order] % nn], axis=1)
    cnt = np.bincount(face_edge, minlength=F)
    if np.any(cnt > 2):
        raise ValueError("an edge is shared by more than two cells")

    # neighbour across each slot
    srt = np.argsort(face_edge, kind="stable")
    e_sorted = face_edge[srt]
    start = np.searchsorted(e_sorted, np.arange(F))
    s0 = srt[start]
    s1 = srt[np.minimum(start + 1, M - 1)]
    pair0 = s0.astype(np.int64)                        # first slot of each edge
    pair1 = np.full(F, -1, dtype=np.int64)             # second slot (internal edges)
    two = cnt == 2
    pair1[two] = s1[two]
    nbr_slot = np.full(M, -1, dtype=np.int64)
    nbr_slot[pair0[two]] = pair1[two]
    nbr_slot[pair1[two]] = pair0[two]

    # ---- geometry -----------------------------------------------------------
    area, centroid = _polygon_geometry(xy, cell_nodes, deg)
    x1, y1 = xy[a, 0], xy[a, 1]
    x2, y2 = xy[b, 0], xy[b, 1]
    dx = x2 - x1
    dy = y2 - y1
    s = np.sqrt(dy * dy + dx * dx)                     # mesh_2D.jl:554-571
    face_nx = dy / s
    face_ny = (-dx) / s
    ex1, ey1 = xy[edge_nodes[:, 0], 0], xy[edge_nodes[:, 0], 1]
    ex2, ey2 = xy[edge_nodes[:, 1], 0], xy[edge_nodes[:, 1], 1]
    enx, eny = ey2 - ey1, -(ex2 - ex1)                 # mesh_2D.jl:637-652
    elen = np.sqrt(enx * enx + eny * eny)
    edge_normal = np.stack([enx / elen, eny / elen], axis=1)
    face_len = elen[face_edge]

    # ---- boundary edges -> boundaries ----------------------------------------
    bnd_edges = np.nonzero(cnt == 1)[0]
    edge_of_key = {int(k): int(r) for k, r in zip(uniq[order][cnt[np.arange(F)] == 1], bnd_edges)}
    assigned = {}
    boundary_faces = {}
    for bid in sorted(nodestrings):
        if bid not in bc_types:
            continue  # no BC line: edges fall to the default wall
        nodes = [int(v) for v in nodestrings[bid]]
        lst = []
        for p, q in zip(nodes[:-1], nodes[1:]):
            k = min(p, q) * nn + max(p, q)
            if k not in edge_of_key:
                raise ValueError(f"boundary edge ({p + 1},{q + 1}) in NodeString {bid} is not a mesh boundary edge")
            e = edge_of_key[k]
            if e in assigned:
                raise ValueError(f"boundary edge ({p + 1},{q + 1}) belongs to two NodeStrings")
            assigned[e] = bid
            lst.append(e)
        boundary_faces[bid] = lst
    types = dict(bc_types)
    unused = [int(e) for e in bnd_edges if int(e) not in assigned]
    if unused:
        default_id = len(boundary_faces) + 1
        boundary_faces[default_id] = unused
        types.setdefault(default_id, "WALL")
    ids = sorted(boundary_faces)
    if ids != list(range(1, len(ids) + 1)):
        raise ValueError(f"boundary ids must be 1..K and contiguous (bc_2D.jl:163-174); got {ids}")

    # ghosts in increasing edge order
    G = bnd_edges.size
    ghost_of_edge = np.full(F, -1, dtype=np.int64)
    ghost_of_edge[bnd_edges] = np.arange(G)
    ghost_edge = bnd_edges
    ghost_slot = pair0[bnd_edges]
    ghost_cell = slot_cell[ghost_slot]
    forward = a[ghost_slot] == edge_nodes[bnd_edges, 0]   # cell walks min -> max
    direction = np.where(forward, 1.0, -1.0)
    ghost_nx = direction * edge_normal[bnd_edges, 0]
    ghost_ny = direction * edge_normal[bnd_edges, 1]
    ghost_len = elen[bnd_edges]
    ghost_kind = np.full(G, -1, dtype=np.int32)
    ghost_bc = np.full(G, -1, dtype=np.int32)
    kinds = {bid: bc_kind(types[bid]) for bid in ids}
    inlet_ids = [bid for bid in ids if kinds[bid] == BC_INLETQ]
    exit_ids = [bid for bid in ids if kinds[bid] == BC_EXITH]
    inlet_lists = []
    for bid in ids:
        g = ghost_of_edge[np.asarray(boundary_faces[bid], dtype=np.int64)]
        ghost_kind[g] = kinds[bid]
        if kinds[bid] == BC_INLETQ:
            ghost_bc[g] = inlet_ids.index(bid)
            inlet_lists.append(g)
        elif kinds[bid] == BC_EXITH:
            ghost_bc[g] = exit_ids.index(bid)
    if np.any(ghost_kind < 0):
        raise ValueError("a boundary edge has no boundary condition")
    inlet_off = np.zeros(len(inlet_lists) + 1, dtype=np.int32)
    if inlet_lists:
        inlet_off[1:] = np.cumsum([len(x) for x in inlet_lists])
        inlet_ghost = np.concatenate(inlet_lists).astype(np.int32)
    else:
        inlet_ghost = np.zeros(0, dtype=np.int32)

    face_nbr = np.where(nbr_slot >= 0, slot_cell[np.maximum(nbr_slot, 0)], 0)
    bslot = nbr_slot < 0
    face_nbr[bslot] = ~ghost_of_edge[face_edge[bslot]]

    by_kind = {}
    for bid in ids:
        by_kind.setdefault(kinds[bid], []).append(bid)

    return Mesh2D(
        node_xy=xy, cell_nodes=np.where(cell_nodes >= 0, cell_nodes, -1).astype(np.int64),
        cell_deg=deg, edge_nodes=edge_nodes,
        face_off=face_off.astype(np.int32), face_nbr=face_nbr.astype(np.int32),
        face_edge=face_edge.astype(np.int32),
        face_nx=np.ascontiguousarray(face_nx), face_ny=np.ascontiguousarray(face_ny),
        face_len=np.ascontiguousarray(face_len), area=area, centroid=centroid,
        ghost_edge=ghost_edge.astype(np.int32), ghost_cell=ghost_cell.astype(np.int32),
        ghost_slot=ghost_slot.astype(np.int32), ghost_kind=ghost_kind, ghost_bc=ghost_bc,
        ghost_nx=np.ascontiguousarray(ghost_nx), ghost_ny=np.ascontiguousarray(ghost_ny),
        ghost_len=np.ascontiguousarray(ghost_len),
        n_inlet=len(inlet_ids), n_exit=len(exit_ids), inlet_off=inlet_off, inlet_ghost=inlet_ghost,
        boundary_ids=by_kind,
    )


def nodes_to_cells(mesh: Mesh2D, z_nodes: np.ndarray) -> np.ndarray:
    """Cell value = mean of its node values, a sequential sum divided by the
    count (fvm_schemes_2D.jl:108-116; SRH_2D_SRHGeom.jl:193-199)."""
    N, maxdeg = mesh.cell_nodes.shape
    acc = z_nodes[mesh.cell_nodes[:, 0]] + z_nodes[mesh.cell_nodes[:, 1]]
    for k in range(2, maxdeg):
        active = k < mesh.cell_deg
        acc = np.where(active, acc + z_nodes[np.maximum(mesh.cell_nodes[:, k], 0)], acc)
    return acc / mesh.cell_deg


def bed_data(mesh: Mesh2D, zb_cells: np.ndarray):
    """zb_ghost, S0x, S0y from zb_cells (process_bed_2D.jl:46-66 with
    fvm_schemes_2D.jl:3-30, 89-105, 133-167): face value = mean of the face's
    two cells (the cell itself on the boundary), S0 = -1.0 * Green-Gauss
    gradient, accumulated over the cell's faces in local order from 0."""
    N = mesh.n_cells
    zb_ghost = zb_cells[mesh.ghost_cell]
    gx = np.zeros(N)
    gy = np.zeros(N)
    off = mesh.face_off.astype(np.int64)
    deg = mesh.cell_deg
    nbr = mesh.face_nbr
    for k in range(int(deg.max())):
        active = k < deg
        s = np.where(active, off[:-1] + k, 0)
        j = nbr[s]
        zf = np.where(j >= 0, (zb_cells + zb_cells[np.maximum(j, 0)]) / 2.0, zb_cells)
        gx = np.where(active, gx + mesh.face_nx[s] * zf * mesh.face_len[s], gx)
        gy = np.where(active, gy + mesh.face_ny[s] * zf * mesh.face_len[s], gy)
    return zb_ghost, -1.0 * (gx / mesh.area), -1.0 * (gy / mesh.area)


def morton_order(xy: np.ndarray, bits: int = 21) -> np.ndarray:
    """Permutation sorting points along a Z-order (Morton) curve."""
    lo = xy.min(axis=0)
    span = np.maximum(xy.max(axis=0) - lo, 1e-300)
    q = np.minimum(((xy - lo) / span.max() * (2**bits - 1)).astype(np.uint64), 2**bits - 1)

    def spread(v):
        v = v & np.uint64(0x1FFFFF)
        v = (v | (v << np.uint64(32))) & np.uint64(0x1F00000000FFFF)
        v = (v | (v << np.uint64(16))) & np.uint64(0x1F0000FF0000FF)
        v = (v | (v << np.uint64(8))) & np.uint64(0x100F00F00F00F00F)
        v = (v | (v << np.uint64(4))) & np.uint64(0x10C30C30C30C30C3)
        v = (v | (v << np.uint64(2))) & np.uint64(0x1249249249249249)
        return v

    code = spread(q[:, 0]) | (spread(q[:, 1]) << np.uint64(1))
    return np.argsort(code, kind="stable")


def concat_meshes(meshes) -> Mesh2D:
    """Block-diagonal union of independent meshes laid end to end: mesh r's
    nodes, cells, edges, slots, ghosts, inlets and exits follow those of
    mesh r-1 (the instance layout of hg_batch_create)."""
    cat = np.concatenate
    nn = np.cumsum([0] + [m.node_xy.shape[0] for m in meshes])
    nc = np.cumsum([0] + [m.n_cells for m in meshes])
    ne = np.cumsum([0] + [m.n_faces for m in meshes])
    ns = np.cumsum([0] + [int(m.face_off[-1]) for m in meshes])
    ng = np.cumsum([0] + [m.n_ghosts for m in meshes])
    ni = np.cumsum([0] + [m.n_inlet for m in meshes])
    nx_ = np.cumsum([0] + [m.n_exit for m in meshes])
    nf = np.cumsum([0] + [int(m.inlet_off[-1]) if m.n_inlet else 0 for m in meshes])
    maxdeg = max(m.cell_nodes.shape[1] for m in meshes)

    def pad(cn, off):
        out = np.full((cn.shape[0], maxdeg), -1, dtype=np.int64)
        out[:, : cn.shape[1]] = np.where(cn >= 0, cn + off, -1)
        return out

    def nbr(m, r):
        j = m.face_nbr.astype(np.int64)
        return np.where(j >= 0, j + nc[r], ~((~j) + ng[r])).astype(np.int32)

    def gbc(m, r):
        b = m.ghost_bc.astype(np.int64)
        b = np.where(m.ghost_kind == BC_INLETQ, b + ni[r], np.where(m.ghost_kind == BC_EXITH, b + nx_[r], b))
        return b.astype(np.int32)

    R = range(len(meshes))
    inlet_off = np.zeros(int(ni[-1]) + 1, dtype=np.int32)
    for r, m in zip(R, meshes):
        if m.n_inlet:
            inlet_off[ni[r] + 1: ni[r + 1] + 1] = np.asarray(m.inlet_off[1:], dtype=np.int64) + nf[r]
    return Mesh2D(
        node_xy=cat([m.node_xy for m in meshes]),
        cell_nodes=cat([pad(m.cell_nodes, nn[r]) for r, m in zip(R, meshes)]),
        cell_deg=cat([m.cell_deg for m in meshes]),
        edge_nodes=cat([m.edge_nodes + nn[r] for r, m in zip(R, meshes)]),
        face_off=cat([[0]] + [np.asarray(m.face_off[1:], dtype=np.int64) + ns[r] for r, m in zip(R, meshes)]).astype(np.int32),
        face_nbr=cat([nbr(m, r) for r, m in zip(R, meshes)]),
        face_edge=cat([np.asarray(m.face_edge, dtype=np.int64) + ne[r] for r, m in zip(R, meshes)]).astype(np.int32),
        face_nx=cat([m.face_nx for m in meshes]), face_ny=cat([m.face_ny for m in meshes]),
        face_len=cat([m.face_len for m in meshes]), area=cat([m.area for m in meshes]),
        centroid=cat([m.centroid for m in meshes]),
        ghost_edge=cat([np.asarray(m.ghost_edge, dtype=np.int64) + ne[r] for r, m in zip(R, meshes)]),
        ghost_cell=cat([m.ghost_cell + nc[r] for r, m in zip(R, meshes)]).astype(np.int32),
        ghost_slot=cat([m.ghost_slot + ns[r] for r, m in zip(R, meshes)]).astype(np.int32),
        ghost_kind=cat([m.ghost_kind for m in meshes]).astype(np.int32),
        ghost_bc=cat([gbc(m, r) for r, m in zip(R, meshes)]),
        ghost_nx=cat([m.ghost_nx for m in meshes]), ghost_ny=cat([m.ghost_ny for m in meshes]),
        ghost_len=cat([m.ghost_len for m in meshes]),
        n_inlet=int(ni[-1]), n_exit=int(nx_[-1]), inlet_off=inlet_off,
        inlet_ghost=cat([np.asarray(m.inlet_ghost, dtype=np.int64) + ng[r] for r, m in zip(R, meshes)
                         if m.n_inlet] or [np.zeros(0, np.int64)]).astype(np.int32),
    )
