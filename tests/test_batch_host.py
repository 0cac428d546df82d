"""Host-side checks of batching: the block-diagonal union of instances is
exactly the instances side by side (oracle runs agree bitwise), and
hg_batch_create rejects an index that leaves its own instance before any
device work."""
import ctypes as C

import numpy as np

from hgswe import capi, cases

import pyoracle


def test_concatenated_case_runs_as_its_instances():
    inst = [cases.channel(1500, seed=1, Lx=800.0, Ly=80.0, Q_in=30.0, nsteps=20, n_bumps=5), cases.dam_break(10, 3),
            cases.channel(1200, seed=2, Lx=600.0, Ly=60.0, Q_in=20.0, nsteps=20, n_bumps=5)]
    B = cases.concat(inst)
    assert B.N == sum(c.N for c in inst) and B.mesh.n_inlet == 2 and B.mesh.n_exit == 2
    QB = pyoracle.Oracle(B).run(B.Q0, B.dt, 25)
    for c, q in zip(inst, cases.split_state(B, QB)):
        assert np.array_equal(q, pyoracle.Oracle(c).run(c.Q0, B.dt, 25))


def _structs(m, keep, face_nbr=None):
    def P(a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        if a.size == 0:
            a = np.zeros(1, dtype=dt)
        keep.append(a)
        return C.c_void_p(a.ctypes.data)

    nb = m.face_nbr if face_nbr is None else face_nbr
    mesh = capi.hg_mesh(m.n_cells, P(m.face_off, np.int32), P(nb, np.int32), P(m.face_nx, np.float64),
                        P(m.face_ny, np.float64), P(m.face_len, np.float64), P(m.area, np.float64), C.c_void_p(0))
    bcs = capi.hg_bcs(m.n_ghosts, P(m.ghost_slot, np.int32), P(m.ghost_kind, np.int32), P(m.ghost_bc, np.int32),
                      m.n_inlet, m.n_exit, P(m.inlet_off, np.int32), P(m.inlet_ghost, np.int32))
    return mesh, bcs


def test_batch_create_rejects_cross_instance_index():
    a, b = cases.dam_break(6, 3), cases.dam_break(5, 3)
    keep = []
    nb = b.mesh.face_nbr.copy()
    s = int(np.nonzero(nb >= 0)[0][0])
    nb[s] = b.N + 2  # would alias into the next instance once offset
    pa = _structs(a.mesh, keep)
    pb = _structs(b.mesh, keep, face_nbr=nb)
    meshes = (capi.hg_mesh * 2)(pa[0], pb[0])
    bcs = (capi.hg_bcs * 2)(pa[1], pb[1])
    h = C.c_void_p()
    rc = capi.lib().hg_batch_create(C.byref(h), 2, meshes, bcs, C.byref(capi.hg_consts(9.81, 1.0, 1e-3)), 0)
    msg = capi.lib().hg_last_error(None).decode()
    assert rc == capi.HG_ERR_MESH and "instance 1" in msg and "face_nbr" in msg
