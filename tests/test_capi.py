"""CPU checks of the C-ABI library: it loads, it exports every entry point
include/hgswe.h declares, and hg_create's host-side validation rejects
malformed meshes with the documented codes before touching the device."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from hgswe import capi, cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "hgswe.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(hg_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    names = declared_functions()
    assert len(names) >= 14
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(capi.EXPORTS) == sorted(names)


def test_julia_shim_binds_only_exported_symbols():
    shim = open(os.path.join(ROOT, "hydrograd.jl_amd", "julia", "HydrogradHIP.jl")).read()
    used = set(re.findall(r":(hg_[a-z0-9_]+)", shim))
    assert used and used <= set(declared_functions()), used - set(declared_functions())


def _raw_create(case, **over):
    m = case.mesh
    arrs = dict(face_off=m.face_off.copy(), face_nbr=m.face_nbr.copy(), face_nx=m.face_nx, face_ny=m.face_ny,
                face_len=m.face_len, area=m.area.copy(), ghost_slot=m.ghost_slot.copy(),
                ghost_kind=m.ghost_kind.copy(), ghost_bc=m.ghost_bc.copy())
    arrs.update(over)
    keep = {k: np.ascontiguousarray(v) for k, v in arrs.items()}
    P = lambda a: C.c_void_p(a.ctypes.data)  # noqa: E731
    mesh = capi.hg_mesh(m.n_cells, P(keep["face_off"]), P(keep["face_nbr"]), P(keep["face_nx"]),
                        P(keep["face_ny"]), P(keep["face_len"]), P(keep["area"]), C.c_void_p(0))
    inl = np.ascontiguousarray(m.inlet_off)
    ing = np.ascontiguousarray(m.inlet_ghost) if m.inlet_ghost.size else np.zeros(1, np.int32)
    bcs = capi.hg_bcs(m.n_ghosts, P(keep["ghost_slot"]), P(keep["ghost_kind"]), P(keep["ghost_bc"]),
                      m.n_inlet, m.n_exit, P(inl), P(ing))
    consts = capi.hg_consts(9.81, 1.0, 1e-3)
    h = C.c_void_p()
    rc = capi.lib().hg_create(C.byref(h), C.byref(mesh), C.byref(bcs), C.byref(consts), 0)
    msg = capi.lib().hg_last_error(None).decode()
    return rc, msg


@pytest.fixture(scope="module")
def small_case():
    return cases.dam_break(nx=6, ny=3)


def test_create_rejects_nonpositive_area(small_case):
    a = small_case.mesh.area.copy()
    a[4] = -a[4]
    rc, msg = _raw_create(small_case, area=a)
    assert rc == capi.HG_ERR_MESH and "area" in msg


def test_create_rejects_asymmetric_neighbours(small_case):
    nb = small_case.mesh.face_nbr.copy()
    s = int(np.nonzero(nb >= 0)[0][0])
    nb[s] = (nb[s] + 5) % small_case.N
    rc, msg = _raw_create(small_case, face_nbr=nb)
    assert rc == capi.HG_ERR_MESH


def test_create_rejects_bad_ghost_map(small_case):
    gs = small_case.mesh.ghost_slot.copy()
    gs[[0, 1]] = gs[[1, 0]]
    rc, msg = _raw_create(small_case, ghost_slot=gs)
    assert rc == capi.HG_ERR_MESH and "ghost_slot" in msg
    gk = small_case.mesh.ghost_kind.copy()
    gk[0] = 7
    rc, _ = _raw_create(small_case, ghost_kind=gk)
    assert rc == capi.HG_ERR_MESH


def test_create_rejects_null_arguments():
    h = C.c_void_p()
    assert capi.lib().hg_create(C.byref(h), None, None, None, 0) == capi.HG_ERR_ARG
