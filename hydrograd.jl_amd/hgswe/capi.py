"""ctypes binding of libhgswe.so (include/hgswe.h).

The product path: every call goes to the HIP library; when the library is
missing or cannot load, the import of this module's ``lib()`` raises --
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libhgswe.so")
_LIB = None

HG_OK, HG_ERR_ARG, HG_ERR_HIP, HG_ERR_MESH, HG_ERR_CONVEYANCE, HG_ERR_NONFINITE, HG_ERR_STATE = range(7)
HG_PARAM_NONE, HG_PARAM_ZB, HG_PARAM_N_ZONE, HG_PARAM_N_CELL, HG_PARAM_Q, HG_PARAM_UDE = range(6)
HG_REORDER = 1
HG_CELL_KERNEL = 2
HG_NLAW_OFF, HG_NLAW_POWER, HG_NLAW_SIGMOID, HG_NLAW_INVERSE, HG_NLAW_H_UMAG_KS = range(5)
HG_LUX_OFF, HG_LUX_MANNING_H, HG_LUX_MANNING_H_UMAG_KS, HG_LUX_FLOW_RESISTANCE = range(4)
# UDE_choice names (process_ManningN_2D.jl:216-273, process_FlowResistance_UDE.jl:3-46)
LUX_CHOICES = {"ManningN_h": HG_LUX_MANNING_H, "ManningN_h_Umag_ks": HG_LUX_MANNING_H_UMAG_KS,
               "FlowResistance": HG_LUX_FLOW_RESISTANCE}
# get_activation names (process_UDE.jl:93-107)
ACTIVATIONS = {"relu": 0, "leakyrelu": 1, "sigmoid": 2, "tanh": 3, "softplus": 4}
# create_manning_function's type names (process_ManningN_2D.jl:122-134)
NLAW_KINDS = {"power_law": HG_NLAW_POWER, "sigmoid": HG_NLAW_SIGMOID, "inverse": HG_NLAW_INVERSE,
              "h_Umag_ks": HG_NLAW_H_UMAG_KS}

# every function declared in include/hgswe.h
EXPORTS = [
    "hg_create", "hg_destroy", "hg_last_error", "hg_set_fields", "hg_set_active_param", "hg_rhs", "hg_step",
    "hg_set_params", "hg_state_upload", "hg_state_download", "hg_advance", "hg_cfl_dt", "hg_time_advance",
    "hg_info", "hg_probe_copy", "hg_rhs_vjp", "hg_adjoint", "hg_grad_probes", "hg_batch_create",
    "hg_set_tile_config", "hg_get_tile_config", "hg_set_adjoint_tile", "hg_last_checkpoint_interval", "hg_adjoint_forward",
    "hg_adjoint_reverse", "hg_autotune", "hg_set_ude", "hg_ude_n", "hg_ude_grad",
    "hg_mfma_probe", "hg_tsit5_solve", "hg_tsit5_tableau", "hg_jvp", "hg_batch_set_dt", "hg_set_n_law",
    "hg_n_law_eval", "hg_cell_n", "hg_lux_n_params", "hg_set_lux_ude", "hg_lux_eval", "hg_lux_grad",
    "hg_advance_timed", "hg_advance_region", "hg_prepare_advance", "hg_cfl_dt_noalloc", "hg_fields_create",
]
UDE_WEIGHTS = 353


class hg_lux_net(C.Structure):
    _fields_ = [("choice", C.c_int32), ("n_hidden", C.c_int32), ("width", C.c_int32 * 4), ("act", C.c_int32 * 4),
                ("output_bounds", C.c_double * 2), ("h_bounds", C.c_double * 2), ("u_bounds", C.c_double * 2),
                ("ks_bounds", C.c_double * 2), ("ln_epsilon", C.c_double)]


def lux_net(choice, hidden, activations, output_bounds, h_bounds, u_bounds=(0.0, 1.0), ks_bounds=(0.0, 1.0),
            ln_epsilon=0.0):
    """hg_lux_net from the reference's UDE_NN_config fields (names or codes)."""
    hidden = list(hidden)
    acts = [ACTIVATIONS.get(a, a) for a in activations]
    if len(hidden) != len(acts) or not 1 <= len(hidden) <= 4:
        raise ValueError("one activation per hidden layer, 1..4 layers")
    w = (C.c_int32 * 4)(*(hidden + [0] * (4 - len(hidden))))
    a = (C.c_int32 * 4)(*(acts + [0] * (4 - len(acts))))
    d2 = lambda b: (C.c_double * 2)(float(b[0]), float(b[1]))
    return hg_lux_net(LUX_CHOICES.get(choice, choice), len(hidden), w, a, d2(output_bounds), d2(h_bounds),
                      d2(u_bounds), d2(ks_bounds), float(ln_epsilon))


def lux_n_params(net):
    return int(lib().hg_lux_n_params(C.byref(net)))


class HGError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libhgswe error {code}: {msg}")
        self.code = code


class hg_mesh(C.Structure):
    _fields_ = [("n_cells", C.c_int64), ("face_off", C.c_void_p), ("face_nbr", C.c_void_p),
                ("face_nx", C.c_void_p), ("face_ny", C.c_void_p), ("face_len", C.c_void_p),
                ("area", C.c_void_p), ("centroid_xy", C.c_void_p)]


class hg_bcs(C.Structure):
    _fields_ = [("n_ghosts", C.c_int64), ("ghost_slot", C.c_void_p), ("ghost_kind", C.c_void_p),
                ("ghost_bc", C.c_void_p), ("n_inlet", C.c_int32), ("n_exit", C.c_int32),
                ("inlet_off", C.c_void_p), ("inlet_ghost", C.c_void_p)]


class hg_consts(C.Structure):
    _fields_ = [("g", C.c_double), ("k_n", C.c_double), ("h_small", C.c_double)]


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() (no CPU fallback exists)")
        L = C.CDLL(LIB_PATH)
        L.hg_last_error.restype = C.c_char_p
        L.hg_last_error.argtypes = [C.c_void_p]
        L.hg_info.restype = C.c_char_p
        L.hg_create.argtypes = [C.POINTER(C.c_void_p), C.POINTER(hg_mesh), C.POINTER(hg_bcs), C.POINTER(hg_consts),
                                C.c_int32]
        L.hg_fields_create.argtypes = [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(hg_mesh), C.POINTER(hg_bcs),
                                       C.POINTER(hg_consts), C.c_int32]
        L.hg_batch_create.argtypes = [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(hg_mesh), C.POINTER(hg_bcs),
                                      C.POINTER(hg_consts), C.c_int32]
        L.hg_set_tile_config.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
        L.hg_batch_set_dt.argtypes = [C.c_void_p, C.c_void_p]
        L.hg_get_tile_config.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.hg_autotune.argtypes = [C.c_void_p, C.c_double, C.c_int64, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                  C.POINTER(C.c_double)]
        L.hg_set_ude.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32]
        L.hg_ude_n.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.hg_ude_grad.argtypes = [C.c_void_p, C.c_void_p]
        L.hg_set_n_law.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]
        L.hg_n_law_eval.argtypes = [C.c_void_p] + [C.c_void_p] * 6
        L.hg_cell_n.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.hg_lux_n_params.argtypes = [C.c_void_p]
        L.hg_lux_n_params.restype = C.c_int64
        L.hg_set_lux_ude.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.hg_lux_eval.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.hg_lux_grad.argtypes = [C.c_void_p, C.c_void_p]
        L.hg_mfma_probe.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.hg_tsit5_solve.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_double, C.c_double,
                                     C.c_int64, C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_void_p]
        L.hg_tsit5_tableau.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.hg_jvp.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_double,
                             C.c_int64, C.c_void_p, C.c_void_p]
        for name in ("hg_destroy",):
            getattr(L, name).argtypes = [C.c_void_p]
        L.hg_set_fields.argtypes = [C.c_void_p] + [C.c_void_p] * 9
        L.hg_set_active_param.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32]
        L.hg_rhs.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]
        L.hg_step.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_double, C.c_int64]
        L.hg_set_params.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
        L.hg_state_upload.argtypes = [C.c_void_p, C.c_void_p]
        L.hg_state_download.argtypes = [C.c_void_p, C.c_void_p]
        L.hg_advance.argtypes = [C.c_void_p, C.c_double, C.c_int64]
        L.hg_advance_timed.argtypes = [C.c_void_p, C.c_double, C.c_int64, C.POINTER(C.c_double)]
        L.hg_advance_region.argtypes = [C.c_void_p, C.c_double, C.c_int64, C.POINTER(C.c_double)]
        L.hg_prepare_advance.argtypes = [C.c_void_p, C.c_double]
        L.hg_cfl_dt.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.POINTER(C.c_double)]
        L.hg_cfl_dt_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.POINTER(C.c_double)]
        L.hg_time_advance.argtypes = [C.c_void_p, C.c_double, C.c_int64, C.c_int32, C.POINTER(C.c_double),
                                      C.POINTER(C.c_double)]
        L.hg_probe_copy.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.POINTER(C.c_double)]
        L.hg_rhs_vjp.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p] + [C.c_void_p] * 4
        L.hg_adjoint.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_double, C.c_int64, C.c_int64,
                                 C.c_void_p] + [C.c_void_p] * 4
        L.hg_grad_probes.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_double, C.c_int64,
                                     C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.POINTER(C.c_double)] + \
            [C.c_void_p] * 3
        _LIB = L
    return _LIB


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else C.c_void_p(0)


class Context:
    """Owns one hg_ctx (one HIP stream) built from a ``cases.Case``.

    A batch (``cases.concat``) of instances is created with
    hg_batch_create from the instances' own meshes; every other call then
    takes the batch's concatenated arrays.  With share_mesh=True, a batch
    whose instances are all the same mesh is created with hg_fields_create
    instead (the mesh stored once, every field advanced in each launch); the
    calls and their arrays are the same.  With fields=False the case's
    fields are not uploaded (set_fields must come before any stepping)."""

    def __init__(self, case, reorder=True, cell_kernel=False, fields=True, share_mesh=False):
        L = lib()
        self.case = case
        self.N = case.mesh.n_cells
        self._keep = []

        def k(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            if a.size == 0:
                a = np.zeros(1, dtype=dt)
            self._keep.append(a)
            return _ptr(a)

        def structs(m):
            mesh = hg_mesh(m.n_cells, k(m.face_off, np.int32), k(m.face_nbr, np.int32), k(m.face_nx, np.float64),
                           k(m.face_ny, np.float64), k(m.face_len, np.float64), k(m.area, np.float64),
                           k(m.centroid.ravel(), np.float64))
            bcs = hg_bcs(m.n_ghosts, k(m.ghost_slot, np.int32), k(m.ghost_kind, np.int32), k(m.ghost_bc, np.int32),
                         m.n_inlet, m.n_exit, k(m.inlet_off, np.int32), k(m.inlet_ghost, np.int32))
            return mesh, bcs

        consts = hg_consts(case.g, case.k_n, case.h_small)
        h = C.c_void_p()
        flags = (HG_REORDER if reorder else 0) | (HG_CELL_KERNEL if cell_kernel else 0)
        inst = case.extra.get("instances") if case.extra else None
        if inst and share_mesh:
            m0 = inst[0].mesh
            if any(c.mesh is not m0 for c in inst):
                raise ValueError("share_mesh needs a batch of one mesh (the same Mesh2D object in every instance)")
            mesh, bcs = structs(m0)
            rc = L.hg_fields_create(C.byref(h), len(inst), C.byref(mesh), C.byref(bcs), C.byref(consts), flags)
        elif inst:
            pairs = [structs(c.mesh) for c in inst]
            meshes = (hg_mesh * len(inst))(*[p[0] for p in pairs])
            bcs = (hg_bcs * len(inst))(*[p[1] for p in pairs])
            rc = L.hg_batch_create(C.byref(h), len(inst), meshes, bcs, C.byref(consts), flags)
        else:
            mesh, bcs = structs(case.mesh)
            rc = L.hg_create(C.byref(h), C.byref(mesh), C.byref(bcs), C.byref(consts), flags)
        if rc:
            raise HGError(rc, L.hg_last_error(None).decode())
        self.h = h
        self._keep = []
        if fields:
            self.set_fields(case)

    def _chk(self, rc):
        if rc:
            raise HGError(rc, lib().hg_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            lib().hg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_fields(self, case, **over):
        vals = dict(hstill=case.hstill, hstill_ghost=case.hstill_ghost, zb=case.zb, zb_ghost=case.zb_ghost,
                    S0x=case.S0x, S0y=case.S0y, n=case.n, inletQ=case.inletQ, exitWSE=case.exitWSE)
        vals.update(over)
        arrs = []
        for name in ("hstill", "hstill_ghost", "zb", "zb_ghost", "S0x", "S0y", "n", "inletQ", "exitWSE"):
            a = np.ascontiguousarray(vals[name], dtype=np.float64)
            arrs.append(a if a.size else np.zeros(1))
        self._chk(lib().hg_set_fields(self.h, *[_ptr(a) for a in arrs]))

    def set_active_param(self, kind, zone=None, n_zones=0):
        z = None if zone is None else np.ascontiguousarray(zone, dtype=np.int32)
        self._chk(lib().hg_set_active_param(self.h, int(kind), _ptr(z), int(n_zones)))

    def rhs(self, Q, params=None):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        out = np.empty(3 * self.N)
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
        self._chk(lib().hg_rhs(self.h, _ptr(Q), _ptr(p), 0 if p is None else p.size, _ptr(out)))
        return out

    def step(self, Q, dt, nsteps, params=None):
        Q = np.array(Q, dtype=np.float64, copy=True)
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
        self._chk(lib().hg_step(self.h, _ptr(Q), _ptr(p), 0 if p is None else p.size, C.c_double(dt),
                                C.c_int64(nsteps)))
        return Q

    def set_params(self, params=None):
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
        self._chk(lib().hg_set_params(self.h, _ptr(p), 0 if p is None else p.size))

    def upload(self, Q):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        self._chk(lib().hg_state_upload(self.h, _ptr(Q)))

    def download(self):
        out = np.empty(3 * self.N)
        self._chk(lib().hg_state_download(self.h, _ptr(out)))
        return out

    def advance(self, dt, nsteps):
        self._chk(lib().hg_advance(self.h, C.c_double(dt), C.c_int64(nsteps)))

    def cfl_dt(self, Q, cfl):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        dt = C.c_double()
        self._chk(lib().hg_cfl_dt(self.h, _ptr(Q), C.c_double(cfl), C.byref(dt)))
        return dt.value

    def cfl_dt_noalloc(self, Q, cfl):
        """cfl_dt through the entry point that allocates nothing (static
        device scratch)."""
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        dt = C.c_double()
        self._chk(lib().hg_cfl_dt_noalloc(self.h, _ptr(Q), C.c_double(cfl), C.byref(dt)))
        return dt.value

    def rhs_vjp(self, Q, ybar, params=None):
        """Cotangents (Q_bar, n_bar, zb_bar, inletQ_bar) of one RHS evaluation."""
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        y = np.ascontiguousarray(ybar, dtype=np.float64)
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
        qb = np.empty(3 * self.N)
        nb = np.empty(self.N)
        zb = np.empty(self.N)
        ib = np.zeros(max(1, self.case.mesh.n_inlet))
        self._chk(lib().hg_rhs_vjp(self.h, _ptr(Q), _ptr(p), 0 if p is None else p.size, _ptr(y), _ptr(qb),
                                   _ptr(nb), _ptr(zb), _ptr(ib)))
        return qb, nb, zb, ib[: self.case.mesh.n_inlet]

    def adjoint(self, Q0, dt, nsteps, kck, QN_bar, params=None):
        Q0 = np.ascontiguousarray(Q0, dtype=np.float64)
        lb = np.ascontiguousarray(QN_bar, dtype=np.float64)
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
        qb = np.empty(3 * self.N)
        nb = np.empty(self.N)
        zb = np.empty(self.N)
        ib = np.zeros(max(1, self.case.mesh.n_inlet))
        self._chk(lib().hg_adjoint(self.h, _ptr(Q0), _ptr(p), 0 if p is None else p.size, C.c_double(dt),
                                   C.c_int64(nsteps), C.c_int64(kck), _ptr(lb), _ptr(qb), _ptr(nb), _ptr(zb),
                                   _ptr(ib)))
        return qb, nb, zb, ib[: self.case.mesh.n_inlet]

    def adjoint_forward(self, Q0, dt, nsteps, kck, params=None):
        """The recorded forward sweep of adjoint(): returns the final state."""
        Q0 = np.ascontiguousarray(Q0, dtype=np.float64)
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
        QN = np.empty(3 * self.N)
        self._chk(lib().hg_adjoint_forward(self.h, _ptr(Q0), _ptr(p), 0 if p is None else p.size, C.c_double(dt),
                                           C.c_int64(nsteps), C.c_int64(kck), _ptr(QN)))
        return QN

    def adjoint_reverse(self, QN_bar, wrt=("n", "zb", "Q")):
        """The reverse sweep after adjoint_forward: (Q0_bar, n_bar, zb_bar,
        inletQ_bar); a cotangent not named in `wrt` is not formed (None)."""
        lb = np.ascontiguousarray(QN_bar, dtype=np.float64)
        qb = np.empty(3 * self.N)
        nb = np.empty(self.N) if "n" in wrt else None
        zb = np.empty(self.N) if "zb" in wrt else None
        ib = np.zeros(max(1, self.case.mesh.n_inlet))
        self._chk(lib().hg_adjoint_reverse(self.h, _ptr(lb), _ptr(qb), _ptr(nb), _ptr(zb), _ptr(ib)))
        return qb, nb, zb, (ib[: self.case.mesh.n_inlet] if "Q" in wrt else None)

    def grad_probes(self, Q0, dt, nsteps, kck, probe, target, params=None, wrt=("n", "zb", "Q")):
        """(loss, d/dn, d/dzb, d/dQ_inlet) of sum_p (WSE_p - target_p)^2 after
        nsteps; a cotangent not named in `wrt` is not formed (None)."""
        Q0 = np.ascontiguousarray(Q0, dtype=np.float64)
        pr = np.ascontiguousarray(probe, dtype=np.int32)
        tg = np.ascontiguousarray(target, dtype=np.float64)
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
        loss = C.c_double()
        nb = np.empty(self.N) if "n" in wrt else None
        zb = np.empty(self.N) if "zb" in wrt else None
        ib = np.zeros(max(1, self.case.mesh.n_inlet))
        self._chk(lib().hg_grad_probes(self.h, _ptr(Q0), _ptr(p), 0 if p is None else p.size, C.c_double(dt),
                                       C.c_int64(nsteps), C.c_int64(kck), C.c_int64(pr.size), _ptr(pr), _ptr(tg),
                                       C.byref(loss), _ptr(nb), _ptr(zb), _ptr(ib)))
        return loss.value, nb, zb, (ib[: self.case.mesh.n_inlet] if "Q" in wrt else None)

    def prepare(self, dt):
        """Instantiates the stepping graph for dt and the dispatch events
        (no steps), so that the next timed advance builds nothing."""
        self._chk(lib().hg_prepare_advance(self.h, C.c_double(dt)))

    def advance_region(self, dt, nsteps):
        """advance() bracketed by HIP events on the context stream: returns
        the GPU time per step in ms."""
        ms = C.c_double()
        self._chk(lib().hg_advance_region(self.h, C.c_double(dt), C.c_int64(nsteps), C.byref(ms)))
        return ms.value

    def advance_timed(self, dt, nsteps):
        """advance() with events on every step kernel's dispatch: returns the
        mean kernel duration in ms."""
        ms = C.c_double()
        self._chk(lib().hg_advance_timed(self.h, C.c_double(dt), C.c_int64(nsteps), C.byref(ms)))
        return ms.value

    def time_advance(self, dt, nsteps, reps):
        a = C.c_double()
        b = C.c_double()
        self._chk(lib().hg_time_advance(self.h, C.c_double(dt), C.c_int64(nsteps), C.c_int32(reps), C.byref(a),
                                        C.byref(b)))
        return a.value, b.value

    def jvp(self, Q0, dt, nsteps, dQ0=None, params=None, dparams=None):
        """Forward-mode derivative of nsteps Euler steps along (dQ0, dparams):
        returns (QN, dQN)."""
        Q0 = np.ascontiguousarray(Q0, dtype=np.float64)
        d0 = None if dQ0 is None else np.ascontiguousarray(dQ0, dtype=np.float64)
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
        dp = None if dparams is None else np.ascontiguousarray(dparams, dtype=np.float64)
        QN = np.empty(3 * self.N)
        dQN = np.empty(3 * self.N)
        self._chk(lib().hg_jvp(self.h, _ptr(Q0), _ptr(d0), _ptr(p), 0 if p is None else p.size, _ptr(dp),
                               C.c_double(dt), C.c_int64(nsteps), _ptr(QN), _ptr(dQN)))
        return QN, dQN

    def tsit5(self, Q0, t0, t1, n_save, abstol=1e-6, reltol=1e-3, dt0=1e-3, params=None):
        """Adaptive Tsit5 from Q0 over [t0, t1]; returns (n_save + 1, 3N) saved
        states and (accepted, rejected, rhs evaluations)."""
        Q0 = np.ascontiguousarray(Q0, dtype=np.float64)
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
        out = np.empty((n_save + 1, 3 * self.N))
        st = np.zeros(3, dtype=np.int64)
        self._chk(lib().hg_tsit5_solve(self.h, _ptr(Q0), _ptr(p), 0 if p is None else p.size, C.c_double(t0),
                                       C.c_double(t1), C.c_int64(n_save), C.c_double(abstol), C.c_double(reltol),
                                       C.c_double(dt0), _ptr(out), _ptr(st)))
        return out, tuple(int(v) for v in st)

    def set_ude(self, weights, set_of_cell=None, precision=32):
        """Manning n from the MLP closure; weights (n_sets, 353) or (353,);
        weights=None turns it off."""
        if weights is None:
            self._chk(lib().hg_set_ude(self.h, 0, C.c_void_p(0), C.c_void_p(0), 32))
            self._ude_sets = 0
            return
        w = np.ascontiguousarray(weights, dtype=np.float64).reshape(-1, UDE_WEIGHTS)
        s = None if set_of_cell is None else np.ascontiguousarray(set_of_cell, dtype=np.int32)
        self._chk(lib().hg_set_ude(self.h, w.shape[0], _ptr(w), _ptr(s), int(precision)))
        self._ude_sets = w.shape[0]

    def ude_n(self, Q):
        n = np.empty(self.N)
        self._chk(lib().hg_ude_n(self.h, _ptr(np.ascontiguousarray(Q, dtype=np.float64)), _ptr(n)))
        return n

    def ude_grad(self):
        """Weight cotangents of the last reverse pass: (353,) for one weight
        set, (n_sets, 353) for several."""
        sets = getattr(self, "_ude_sets", 0)
        if sets < 1:
            raise RuntimeError("hg_ude_grad: the UDE closure is off")
        g = np.empty((sets, UDE_WEIGHTS))
        self._chk(lib().hg_ude_grad(self.h, _ptr(g)))
        return g[0] if sets == 1 else g

    def set_n_law(self, kind, n_lower=0.0, n_upper=0.0, k=1.0, h_mid=1.0, ks=None):
        """Variable Manning n (the forward simulation's "variable" option):
        kind is HG_NLAW_* or a reference type name ("power_law", "sigmoid",
        "inverse", "h_Umag_ks"); ks per cell for h_Umag_ks.  kind=None or
        HG_NLAW_OFF turns it off."""
        kind = HG_NLAW_OFF if kind is None else NLAW_KINDS.get(kind, kind)
        p = np.array([n_lower, n_upper, k, h_mid], dtype=np.float64)
        ks = None if ks is None else np.ascontiguousarray(ks, dtype=np.float64)
        if ks is not None and ks.size != self.N:
            raise ValueError("ks needs one value per cell")
        self._chk(lib().hg_set_n_law(self.h, int(kind), _ptr(p), _ptr(ks)))

    def n_law_eval(self, h, Umag):
        """(n, h/ks, f, Re) of the law at the given depths and velocity magnitudes."""
        h = np.ascontiguousarray(h, dtype=np.float64)
        U = np.ascontiguousarray(Umag, dtype=np.float64)
        if h.size != self.N or U.size != self.N:
            raise ValueError("h and Umag need one value per cell")
        outs = [np.empty(self.N) for _ in range(4)]
        self._chk(lib().hg_n_law_eval(self.h, _ptr(h), _ptr(U), *[_ptr(o) for o in outs]))
        return tuple(outs)

    def cell_n(self, Q):
        """Manning n per cell that the RHS uses at state Q."""
        n = np.empty(self.N)
        self._chk(lib().hg_cell_n(self.h, _ptr(np.ascontiguousarray(Q, dtype=np.float64)), _ptr(n)))
        return n

    def set_lux_ude(self, net, params=None, ks=None):
        """The reference's UDE network (hg_lux_net from lux_net(...)) with its
        flat weights; net=None turns it off."""
        if net is None:
            self._chk(lib().hg_set_lux_ude(self.h, C.c_void_p(0), C.c_void_p(0), C.c_void_p(0)))
            self.lux_P = 0
            return
        P = lux_n_params(net)
        w = np.ascontiguousarray(params, dtype=np.float64)
        if P < 0 or w.size != P:
            raise ValueError(f"the network takes {P} weights, got {w.size}")
        ks = None if ks is None else np.ascontiguousarray(ks, dtype=np.float64)
        if ks is not None and ks.size != self.N:
            raise ValueError("ks needs one value per cell")
        self._net = net
        self._chk(lib().hg_set_lux_ude(self.h, C.byref(net), _ptr(w), _ptr(ks)))
        self.lux_P = P

    def lux_eval(self, Q):
        """The network's output per cell (n, or the friction magnitude) at state Q."""
        out = np.empty(self.N)
        self._chk(lib().hg_lux_eval(self.h, _ptr(np.ascontiguousarray(Q, dtype=np.float64)), _ptr(out)))
        return out

    def lux_grad(self):
        g = np.empty(self.lux_P)
        self._chk(lib().hg_lux_grad(self.h, _ptr(g)))
        return g

    def batch_set_dt(self, dts):
        """Per-instance time steps of a batch (None: back to the scalar dt)."""
        d = None if dts is None else np.ascontiguousarray(dts, dtype=np.float64)
        self._chk(lib().hg_batch_set_dt(self.h, _ptr(d)))

    def set_tile_config(self, tile_cells, xcd_chunk=0):
        self._chk(lib().hg_set_tile_config(self.h, int(tile_cells), int(xcd_chunk)))

    def last_checkpoint_interval(self):
        """Checkpoint interval of the last gradient call (what kck <= 0 chose)."""
        k = C.c_int64()
        self._chk(lib().hg_last_checkpoint_interval(self.h, C.byref(k)))
        return k.value

    def set_adjoint_tile(self, tile_cells):
        """Tile size of the reverse step (0: the library's choice)."""
        self._chk(lib().hg_set_adjoint_tile(self.h, int(tile_cells)))

    def tile_config(self):
        t, x = C.c_int32(), C.c_int32()
        self._chk(lib().hg_get_tile_config(self.h, C.byref(t), C.byref(x)))
        return t.value, x.value

    def autotune(self, dt, nsteps=64):
        """Times every launch shape from the device state (restored
        afterwards) and keeps the fastest: (tile_cells, xcd_chunk, ms/step)."""
        t, x, ms = C.c_int32(), C.c_int32(), C.c_double()
        self._chk(lib().hg_autotune(self.h, C.c_double(dt), C.c_int64(nsteps), C.byref(t), C.byref(x), C.byref(ms)))
        return t.value, x.value, ms.value


def mfma_probe(precision, A, B):
    """D = A (16x4) B (4x16) on the matrix cores (lane-layout self-check)."""
    A = np.ascontiguousarray(A, dtype=np.float64).reshape(16, 4)
    B = np.ascontiguousarray(B, dtype=np.float64).reshape(4, 16)
    D = np.empty((16, 16))
    rc = lib().hg_mfma_probe(int(precision), _ptr(A), _ptr(B), _ptr(D))
    if rc:
        raise HGError(rc, "hg_mfma_probe failed")
    return D


def tsit5_tableau():
    """(c[7], a[7][6], btilde[7]) of the library's Tsit5."""
    c, a, bt = np.empty(7), np.empty((7, 6)), np.empty(7)
    lib().hg_tsit5_tableau(_ptr(c), _ptr(a), _ptr(bt))
    return c, a, bt


def probe_copy(ctx, nbytes, reps=1):
    g = C.c_double()
    ctx._chk(lib().hg_probe_copy(ctx.h, C.c_int64(nbytes), C.c_int32(reps), C.byref(g)))
    return g.value


def info():
    return lib().hg_info().decode()
