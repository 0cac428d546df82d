"""The oracle's CFL time step pinned by closed forms (CPU).

The time step has no reference counterpart (the reference keeps a CFL field
it never reads and steps with the fixed dt it is given,
custom_ODE_solvers.jl:39); the definition the build uses is

    dt = CFL * min over wet cells (h > h_small) of (2 A / P) / (|u| + sqrt(g h)),

with h = xi + h_still and u = q / h.  On meshes whose cells are all alike the
characteristic length 2A/P has a closed form: dx dy / (dx + dy) for a dx x dy
rectangle and the inradius s / (2 sqrt 3) for an equilateral triangle of side
s.  The cases below are chosen so that each plausible mistake fails one of
them: the perimeter or A/P in place of 2A/P, |u| without the gravity-wave
speed (or the reverse), the depth without h_still, velocities from the
unclamped sign, dry cells taking part in the minimum, or a minimum taken over
the wrong cell.
"""
import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle  # noqa: E402
from hgswe import cases  # noqa: E402
from hgswe import mesh as hm  # noqa: E402

G = 9.81


def _rect_case(nx, ny, Lx, Ly, hstill):
    xy, cells, sides = cases.rect_quads(Lx, Ly, nx, ny)
    ns = {1: sides["bottom"], 2: sides["right"], 3: sides["top"], 4: sides["left"]}
    mesh = hm.build_mesh(xy, cells, ns, {1: "WALL", 2: "WALL", 3: "WALL", 4: "WALL"})
    N = mesh.n_cells
    zb = np.full(N, -hstill)  # wstill = 0 => h_still = -zb
    return cases._finish("rect", mesh, zb, np.full(N, 0.03), np.zeros(N), np.zeros(N), np.zeros(N),
                         np.zeros(N), [], [], 0.01, 1)


def _equilateral_case(nx, ny, s, hstill):
    """Rows of equilateral triangles of side s (odd rows shifted by s/2)."""
    ht = s * math.sqrt(3.0) / 2.0
    nid = lambda i, j: j * (nx + 1) + i  # noqa: E731
    xy = np.array([[i * s + (j % 2) * s / 2.0, j * ht] for j in range(ny + 1) for i in range(nx + 1)])
    tris = []
    for j in range(ny):
        for i in range(nx):
            if j % 2 == 0:
                tris.append([nid(i, j), nid(i + 1, j), nid(i, j + 1)])
                tris.append([nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)])
            else:
                tris.append([nid(i, j), nid(i + 1, j), nid(i + 1, j + 1)])
                tris.append([nid(i, j), nid(i + 1, j + 1), nid(i, j + 1)])
    mesh = hm.build_mesh(xy, np.asarray(tris, dtype=np.int64), {}, {})
    N = mesh.n_cells
    zb = np.full(N, -hstill)
    return cases._finish("equilateral", mesh, zb, np.full(N, 0.03), np.zeros(N), np.zeros(N), np.zeros(N),
                         np.zeros(N), [], [], 0.01, 1)


def _state(c, xi, u, v):
    N = c.N
    h = xi + c.hstill
    return np.concatenate([xi, u * h, v * h])


def _oracle_dt(c, Q, cfl):
    return pyoracle.Oracle(c).cfl_dt(Q, cfl)


def test_rectangles_closed_form():
    dx, dy, hs = 3.0, 2.0, 1.25
    c = _rect_case(10, 4, 10 * dx, 4 * dy, hs)
    assert np.allclose(c.mesh.area, dx * dy)
    N = c.N
    xi = np.full(N, 0.25)  # h = 1.5
    u, v = np.full(N, 0.7), np.full(N, -0.4)
    Q = _state(c, xi, u, v)
    expect = 0.5 * (dx * dy / (dx + dy)) / (math.hypot(0.7, 0.4) + math.sqrt(G * 1.5))
    got = _oracle_dt(c, Q, 0.5)
    assert got == pytest.approx(expect, rel=1e-14)
    # the same cells at rest: only the gravity wave speed remains
    Q0 = _state(c, xi, np.zeros(N), np.zeros(N))
    assert _oracle_dt(c, Q0, 1.0) == pytest.approx(1.2 / math.sqrt(G * 1.5), rel=1e-14)
    # the step scales linearly with the CFL number
    assert _oracle_dt(c, Q, 0.9) == pytest.approx(got * 0.9 / 0.5, rel=1e-14)


def test_equilateral_triangles_inradius():
    s, hs = 4.0, 2.0
    c = _equilateral_case(8, 6, s, hs)
    assert np.all(np.diff(c.mesh.face_off) == 3)
    assert np.allclose(c.mesh.area, math.sqrt(3.0) / 4.0 * s * s)
    N = c.N
    xi = np.zeros(N)  # h = h_still = 2
    Q = _state(c, xi, np.full(N, -1.1), np.full(N, 0.3))
    r = s / (2.0 * math.sqrt(3.0))
    expect = 0.4 * r / (math.hypot(1.1, 0.3) + math.sqrt(G * 2.0))
    assert _oracle_dt(c, Q, 0.4) == pytest.approx(expect, rel=1e-14)


def test_minimum_is_over_wet_cells_only():
    dx, dy, hs = 3.0, 2.0, 1.25
    c = _rect_case(10, 4, 10 * dx, 4 * dy, hs)
    N = c.N
    xi = np.full(N, 0.25)
    u, v = np.full(N, 0.7), np.full(N, -0.4)
    base = _oracle_dt(c, _state(c, xi, u, v), 0.5)
    # dry cells (h = h_small / 2 and h = h_small exactly) with very large
    # discharges would give a far smaller dt if they took part
    xi_d = xi.copy()
    xi_d[[3, 17]] = 0.5e-3 - hs
    xi_d[25] = 1e-3 - hs
    Q = _state(c, xi_d, u, v)
    Q[N + 3] = Q[N + 17] = Q[N + 25] = 50.0
    assert _oracle_dt(c, Q, 0.5) == base
    # one fast wet cell sets the step (the minimum is over cells)
    u2 = u.copy()
    u2[11] = -6.0
    expect = 0.5 * 1.2 / (math.hypot(6.0, 0.4) + math.sqrt(G * 1.5))
    assert _oracle_dt(c, _state(c, xi, u2, v), 0.5) == pytest.approx(expect, rel=1e-14)
    # one shallow wet cell: its slower gravity wave raises its own dt, so the
    # minimum stays with the other cells; one deep cell lowers it
    xi_s = xi.copy()
    xi_s[5] = 0.01 - hs
    assert _oracle_dt(c, _state(c, xi_s, u, v), 0.5) == pytest.approx(base, rel=1e-14)
    xi_s[5] = 20.0 - hs
    expect = 0.5 * 1.2 / (math.hypot(0.7, 0.4) + math.sqrt(G * 20.0))
    assert _oracle_dt(c, _state(c, xi_s, u, v), 0.5) == pytest.approx(expect, rel=1e-14)


def test_numpy_statement_agrees_with_the_oracle():
    """cases.cfl_dt (the host-side statement the generators use) on a
    jittered triangle mesh with a random wet/dry state."""
    c = cases.channel(3_000, seed=3, Lx=600.0, Ly=60.0, nsteps=1)
    rng = np.random.default_rng(1)
    N = c.N
    xi = rng.uniform(-0.2, 0.2, N)
    dry = rng.choice(N, 40, replace=False)
    xi[dry] = 1e-4 - c.hstill[dry]
    Q = np.concatenate([xi, rng.normal(0, 1, N), rng.normal(0, 1, N)])
    assert cases.cfl_dt(c, 0.5, Q) == _oracle_dt(c, Q, 0.5)
