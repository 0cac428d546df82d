"""TEST INFRASTRUCTURE ONLY: ctypes wrapper over oracle/liborc.so, the CPU
restatement of Hydrograd.jl's SWE RHS and Euler step (hg_oracle.hpp).

Only the test-suite, the smoke check and the CPU-baseline leg of bench.py
may import this module; the product path never does.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

D = np.float64
I32 = np.int32


class _Mesh(C.Structure):
    _fields_ = [
        ("N", C.c_int64), ("G", C.c_int64),
        ("face_off", C.c_void_p), ("face_nbr", C.c_void_p),
        ("face_nx", C.c_void_p), ("face_ny", C.c_void_p), ("face_len", C.c_void_p), ("area", C.c_void_p),
        ("ghost_cell", C.c_void_p), ("ghost_kind", C.c_void_p), ("ghost_bc", C.c_void_p),
        ("ghost_nx", C.c_void_p), ("ghost_ny", C.c_void_p), ("ghost_len", C.c_void_p),
        ("n_inlet", C.c_int32), ("n_exit", C.c_int32),
        ("inlet_off", C.c_void_p), ("inlet_ghost", C.c_void_p),
    ]


class _Consts(C.Structure):
    _fields_ = [("g", C.c_double), ("k_n", C.c_double), ("h_small", C.c_double)]


class _Fields(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in
                ("hstill", "hstill_ghost", "zb", "zb_ghost", "S0x", "S0y", "n", "inletQ", "exitWSE", "fmag")]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liborc.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liborc.so is not built (run __graft_entry__.build())")
        L = C.CDLL(path)
        L.orc_last_error.restype = C.c_char_p
        L.orc_cfl_dt.restype = C.c_double
        _LIB = L
    return _LIB


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None and a.size else C.c_void_p(0)


class Oracle:
    """Holds contiguous copies of a Case's arrays and calls the oracle."""

    def __init__(self, case):
        m = case.mesh
        self.case = case
        self.keep = {}

        def k(name, arr, dt):
            a = np.ascontiguousarray(arr, dtype=dt)
            if a.size == 0:
                a = np.zeros(1, dtype=dt)
            self.keep[name] = a
            return C.c_void_p(a.ctypes.data)

        self.mesh = _Mesh(
            m.n_cells, m.n_ghosts,
            k("face_off", m.face_off, I32), k("face_nbr", m.face_nbr, I32),
            k("face_nx", m.face_nx, D), k("face_ny", m.face_ny, D), k("face_len", m.face_len, D),
            k("area", m.area, D),
            k("ghost_cell", m.ghost_cell, I32), k("ghost_kind", m.ghost_kind, I32), k("ghost_bc", m.ghost_bc, I32),
            k("ghost_nx", m.ghost_nx, D), k("ghost_ny", m.ghost_ny, D), k("ghost_len", m.ghost_len, D),
            m.n_inlet, m.n_exit, k("inlet_off", m.inlet_off, I32), k("inlet_ghost", m.inlet_ghost, I32))
        self.consts = _Consts(case.g, case.k_n, case.h_small)
        self.set_fields()

    def set_fields(self, **over):
        c = self.case
        vals = dict(hstill=c.hstill, hstill_ghost=c.hstill_ghost, zb=c.zb, zb_ghost=c.zb_ghost, S0x=c.S0x,
                    S0y=c.S0y, n=c.n, inletQ=c.inletQ, exitWSE=c.exitWSE)
        vals.update(over)
        ptrs = []
        for name in ("hstill", "hstill_ghost", "zb", "zb_ghost", "S0x", "S0y", "n", "inletQ", "exitWSE"):
            a = np.ascontiguousarray(vals[name], dtype=D)
            if a.size == 0:
                a = np.zeros(1)
            self.keep["f_" + name] = a
            ptrs.append(C.c_void_p(a.ctypes.data))
        # fmag: FlowResistance friction magnitudes per cell (None: Manning friction)
        fm = vals.get("fmag")
        if fm is not None:
            fm = np.ascontiguousarray(fm, dtype=D)
            self.keep["f_fmag"] = fm
        ptrs.append(C.c_void_p(0) if fm is None else C.c_void_p(fm.ctypes.data))
        self.fields = _Fields(*ptrs)

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(lib().orc_last_error().decode())

    def rhs(self, Q, recompute_bed=False, return_total_A=False):
        Q = np.ascontiguousarray(Q, dtype=D)
        out = np.empty(3 * self.case.N)
        tA = np.zeros(max(1, self.case.mesh.n_inlet))
        self._check(lib().orc_rhs(C.byref(self.mesh), C.byref(self.consts), C.byref(self.fields),
                                  int(recompute_bed), _p(Q), _p(out), _p(tA)))
        return (out, tA[: self.case.mesh.n_inlet]) if return_total_A else out

    def run(self, Q, dt, nsteps, recompute_bed=False):
        Q = np.array(Q, dtype=D, copy=True)
        self._check(lib().orc_run(C.byref(self.mesh), C.byref(self.consts), C.byref(self.fields),
                                  int(recompute_bed), _p(Q), C.c_double(dt), C.c_int64(nsteps)))
        return Q

    def bed_data(self, zb):
        zb = np.ascontiguousarray(zb, dtype=D)
        zg = np.empty(max(1, self.case.G))
        sx = np.empty(self.case.N)
        sy = np.empty(self.case.N)
        self._check(lib().orc_bed_data(C.byref(self.mesh), _p(zb), _p(zg), _p(sx), _p(sy)))
        return zg[: self.case.G], sx, sy

    def cfl_dt(self, Q, cfl):
        Q = np.ascontiguousarray(Q, dtype=D)
        hs = np.ascontiguousarray(self.case.hstill, dtype=D)
        return lib().orc_cfl_dt(C.byref(self.mesh), C.byref(self.consts), _p(hs), _p(Q), C.c_double(cfl))

    def rhs_jvp(self, Q, dQ=None, dn=None, dzb=None, dQin=None):
        N = self.case.N
        Q = np.ascontiguousarray(Q, dtype=D)
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=D) for a in (dQ, dn, dzb, dQin)]
        out = np.empty(3 * N)
        dout = np.empty(3 * N)
        self._check(lib().orc_rhs_jvp(C.byref(self.mesh), C.byref(self.consts), C.byref(self.fields), _p(Q),
                                      *[(_p(a) if a is not None else C.c_void_p(0)) for a in arrs], _p(out),
                                      _p(dout)))
        return out, dout

    def jvp_run(self, Q0, dt, nsteps, dn=None, dzb=None, dQin=None, probe=None, target=None):
        N = self.case.N
        Q0 = np.ascontiguousarray(Q0, dtype=D)
        probe = np.zeros(0, I32) if probe is None else np.ascontiguousarray(probe, dtype=I32)
        target = np.zeros(0) if target is None else np.ascontiguousarray(target, dtype=D)
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=D) for a in (dn, dzb, dQin)]
        Qout = np.empty(3 * N)
        dQout = np.empty(3 * N)
        loss = C.c_double()
        dloss = C.c_double()
        self._check(lib().orc_jvp_run(
            C.byref(self.mesh), C.byref(self.consts), C.byref(self.fields), _p(Q0), C.c_double(dt),
            C.c_int64(nsteps), *[(_p(a) if a is not None else C.c_void_p(0)) for a in arrs],
            C.c_int64(probe.size), _p(probe), _p(target), _p(Qout), _p(dQout), C.byref(loss), C.byref(dloss)))
        return Qout, dQout, loss.value, dloss.value


    def grad_run(self, Q0, dt, nsteps, probe, target, wrt=("n", "Q", "Q0")):
        """Reverse-mode (tape) gradient of the final-state probe loss
        sum_p (xi_p + hstill_p + zb_p - target_p)^2 after nsteps Euler steps:
        (loss, d/dn, d/dQ_inlet, d/dQ0) and, when "zb" is in `wrt`, d/dzb
        last (the bed data then rebuilt from zb, the active-zb path); a
        gradient not named in `wrt` is None.  One thread."""
        N = self.case.N
        Q0 = np.ascontiguousarray(Q0, dtype=D)
        probe = np.ascontiguousarray(probe, dtype=I32)
        target = np.ascontiguousarray(target, dtype=D)
        gn = np.empty(N) if "n" in wrt else None
        gq = np.empty(max(1, self.case.mesh.n_inlet)) if "Q" in wrt else None
        g0 = np.empty(3 * N) if "Q0" in wrt else None
        gz = np.empty(N) if "zb" in wrt else None
        loss = C.c_double()
        self._check(lib().orc_grad_run(
            C.byref(self.mesh), C.byref(self.consts), C.byref(self.fields), _p(Q0), C.c_double(dt),
            C.c_int64(nsteps), C.c_int64(probe.size), _p(probe), _p(target), C.byref(loss),
            *[(_p(a) if a is not None else C.c_void_p(0)) for a in (gn, gq, g0, gz)]))
        out = (loss.value, gn, (None if gq is None else gq[: self.case.mesh.n_inlet]), g0)
        return out + (gz,) if "zb" in wrt else out

    @staticmethod
    def release_tape():
        """Frees the tape grad_run keeps between calls."""
        lib().orc_release_tape()

    def jvp_state(self, Q0, dQ0, dt, nsteps):
        """Dual run of nsteps Euler steps from (Q0, dQ0): (QN, dQN)."""
        N = self.case.N
        Q0 = np.ascontiguousarray(Q0, dtype=D)
        dQ0 = np.ascontiguousarray(dQ0, dtype=D)
        Qout = np.empty(3 * N)
        dQout = np.empty(3 * N)
        self._check(lib().orc_jvp_state(C.byref(self.mesh), C.byref(self.consts), C.byref(self.fields), _p(Q0),
                                        _p(dQ0), C.c_double(dt), C.c_int64(nsteps), _p(Qout), _p(dQout)))
        return Qout, dQout


def friction(h, qx, qy, n, g=9.81, k_n=1.0, h_small=1e-3):
    arrs = [np.ascontiguousarray(a, dtype=D) for a in (h, qx, qy, n)]
    fx = np.empty_like(arrs[0])
    fy = np.empty_like(arrs[0])
    c = _Consts(g, k_n, h_small)
    rc = lib().orc_friction(C.byref(c), C.c_int64(arrs[0].size), *[_p(a) for a in arrs], _p(fx), _p(fy))
    if rc:
        raise RuntimeError(lib().orc_last_error().decode())
    return fx, fy


def roe(states, normals, g=9.81, hmin=1e-3):
    s = np.ascontiguousarray(states, dtype=D).reshape(-1, 12)
    nrm = np.ascontiguousarray(normals, dtype=D).reshape(-1, 2)
    F = np.empty((s.shape[0], 3))
    rc = lib().orc_roe(C.c_int64(s.shape[0]), _p(s), _p(nrm), C.c_double(g), C.c_double(hmin), _p(F))
    if rc:
        raise RuntimeError(lib().orc_last_error().decode())
    return F


def set_threads(n):
    lib().orc_set_threads(int(n))
