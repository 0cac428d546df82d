"""Parameter inversion and UDE training over the device path: the
reference's inversion loss on the final state (compute_loss_inversion,
swe_2D_inversion.jl:247-483) and its UDE loss (compute_loss_UDE,
swe_2D_UDE.jl:522-599: the same WSE and uv terms), their gradients through
the library's adjoint (hg_adjoint; hg_lux_grad for the network weights), and
the Adam loop the reference drives through Optimization.jl
(swe_2D_inversion.jl:600-732, swe_2D_UDE.jl:360-420; Optimisers.jl Adam).

The forward model is the custom Euler driver (custom_ODE_solve) the adjoint
differentiates; the loss terms and their normalisations follow the
reference:
  WSE  : sum(((WSE - WSE_truth) / (range(WSE_truth) + eps))^2) / N
  uv   : (sum(((u - u_t) / s)^2) + sum(((v - v_t) / s)^2)) / N,
         s = max(max |(u_t, v_t)|, eps(Float64)), u = q_x / (h + eps)
  bound: sum((max(0, lo - p) / (hi - lo))^2 + (max(0, p - hi) / (hi - lo))^2) / N
  slope: bound loss of |grad zb| (Green-Gauss, +eps(Float64) inside the
         root) against [0, slope_upper], / N (zb active only)
with eps = sqrt(eps(Float64)).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import capi

EPS64 = np.finfo(np.float64).eps
EPS = np.sqrt(EPS64)


@dataclass
class LossSettings:
    wse: bool = True
    uv: bool = True
    bound: bool = True
    slope: bool = False
    lower: float = 0.01
    upper: float = 0.06
    slope_upper: float = 0.2


def bound_loss(p, lo, hi):
    """compute_bound_loss (swe_2D_inversion.jl:738-749) and its gradient."""
    p = np.asarray(p, dtype=np.float64)
    r = hi - lo
    lv = np.maximum(0.0, lo - p) / r
    uv = np.maximum(0.0, p - hi) / r
    loss = float(np.sum(lv * lv) + np.sum(uv * uv))
    grad = -2.0 * lv / r + 2.0 * uv / r
    return loss, grad


def gradient_operator(mesh):
    """Green-Gauss cell gradient of a cell field as a sparse (2N x N) matrix:
    grad_i = (1/A_i) sum_k n_ik L_ik z_f, z_f = (z_i + z_j)/2 on internal
    faces and z_i on boundary faces (compute_scalar_gradients_from_cells,
    fvm_schemes_2D.jl:89-167)."""
    from scipy import sparse

    N = mesh.n_cells
    deg = np.diff(mesh.face_off)
    cell = np.repeat(np.arange(N), deg)
    j = mesh.face_nbr.astype(np.int64)
    w = mesh.face_len / mesh.area[cell]
    internal = j >= 0
    rows, cols, vals = [], [], []
    for comp, nrm in ((0, mesh.face_nx), (1, mesh.face_ny)):
        c = nrm * w
        rows += [comp * N + cell[internal], comp * N + cell[internal], comp * N + cell[~internal]]
        cols += [cell[internal], j[internal], cell[~internal]]
        vals += [0.5 * c[internal], 0.5 * c[internal], c[~internal]]
    return sparse.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(2 * N, N))


def slope_loss(G, zb, slope_upper):
    """calc_slope_loss (swe_2D_inversion.jl:752-772) and its gradient w.r.t. zb."""
    N = zb.size
    g = G @ zb
    gx, gy = g[:N], g[N:]
    mag = np.sqrt(gx * gx + gy * gy + EPS64)
    loss, dmag = bound_loss(mag, 0.0, slope_upper)
    dg = np.concatenate([dmag * gx / mag, dmag * gy / mag])
    return loss / N, (G.T @ dg) / N


def final_state_loss(case, QN, truth, settings: LossSettings, params=None, zb=None, G=None):
    """The reference's loss terms on the final state QN (3N), and d(loss)/dQN.

    truth: dict with WSE_truth, u_truth, v_truth (per cell).  params: the
    active parameter vector (for the bound loss).  zb: the bed the WSE uses
    (the active zb when inverting for it).  Returns (total, parts, QN_bar,
    params_bar_direct, zb_bar_direct): the bound / slope terms' own gradient
    w.r.t. params, and the WSE term's direct gradient w.r.t. zb."""
    N = case.N
    zb = case.zb if zb is None else zb
    xi, qx, qy = QN[:N], QN[N:2 * N], QN[2 * N:]
    h = xi + case.hstill
    QN_bar = np.zeros(3 * N)
    zb_bar = np.zeros(N)
    parts = {"wse": 0.0, "uv": 0.0, "bound": 0.0, "slope": 0.0}
    if settings.wse:
        wt = np.asarray(truth["WSE_truth"], dtype=np.float64)
        rng = wt.max() - wt.min()
        lw = (h + zb - wt) / (rng + EPS)
        parts["wse"] = float(np.sum(lw * lw)) / N
        QN_bar[:N] += 2.0 * lw / (rng + EPS) / N
        zb_bar += 2.0 * lw / (rng + EPS) / N
    if settings.uv:
        ut = np.asarray(truth["u_truth"], dtype=np.float64)
        vt = np.asarray(truth["v_truth"], dtype=np.float64)
        vs = max(float(np.max(np.sqrt(ut * ut + vt * vt))), EPS64)
        u, v = qx / (h + EPS), qy / (h + EPS)
        lu, lv = (u - ut) / vs, (v - vt) / vs
        parts["uv"] = (float(np.sum(lu * lu)) + float(np.sum(lv * lv))) / N
        du, dv = 2.0 * lu / vs / N, 2.0 * lv / vs / N
        QN_bar[N:2 * N] += du / (h + EPS)
        QN_bar[2 * N:] += dv / (h + EPS)
        QN_bar[:N] += -(du * u + dv * v) / (h + EPS)
    pb = None if params is None else np.zeros(np.size(params))
    if settings.bound and params is not None:
        lb, gb = bound_loss(params, settings.lower, settings.upper)
        parts["bound"] = lb / N
        pb += gb / N
    if settings.slope and G is not None and params is not None and np.size(params) == N:
        ls, gs = slope_loss(G, np.asarray(params, dtype=np.float64), settings.slope_upper)
        parts["slope"] = ls
        pb += gs
    total = parts["wse"] + parts["uv"] + parts["bound"] + parts["slope"]
    return total, parts, QN_bar, pb, zb_bar


@dataclass
class Adam:
    """Optimisers.jl Adam: m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2,
    p -= lr (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps)."""
    lr: float = 1e-3
    b1: float = 0.9
    b2: float = 0.999
    eps: float = 1e-8
    t: int = 0
    m: np.ndarray = None
    v: np.ndarray = None

    def step(self, p, g):
        if self.m is None:
            self.m = np.zeros_like(p)
            self.v = np.zeros_like(p)
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        return p - self.lr * mh / (np.sqrt(vh) + self.eps)


@dataclass
class InversionResult:
    params: np.ndarray
    loss: list = field(default_factory=list)
    parts: list = field(default_factory=list)
    history: list = field(default_factory=list)


def loss_and_grad(ctx, case, truth, params, active, dt, nsteps, kck, settings: LossSettings, zone=None, G=None):
    """Loss of the final state after nsteps Euler steps and its gradient
    w.r.t. the active parameter vector (zone n, per-cell n, zb or inlet Q)."""
    QN = ctx.step(case.Q0, dt, nsteps, params=params)
    zb = params if active == capi.HG_PARAM_ZB else case.zb
    total, parts, QN_bar, pb, zbd = final_state_loss(case, QN, truth, settings, params=params, zb=zb, G=G)
    _, nb, zbb, qb = ctx.adjoint(case.Q0, dt, nsteps, kck, QN_bar, params=params)
    if active == capi.HG_PARAM_N_ZONE:
        g = np.bincount(zone, weights=nb, minlength=np.size(params))
    elif active == capi.HG_PARAM_N_CELL:
        g = nb.copy()
    elif active == capi.HG_PARAM_ZB:
        g = zbb + zbd  # WSE = h + zb: the final-state loss also reads zb directly
    elif active == capi.HG_PARAM_Q:
        g = qb.copy()
    else:
        raise ValueError("no active parameter")
    return total, parts, g + pb


def invert(ctx, case, truth, params0, active, dt, nsteps, kck=50, lr=1e-3, iters=100, settings=None, zone=None,
           callback=None):
    """Adam on the reference's inversion loss; returns the parameter and
    loss history (the callback bookkeeping of swe_2D_inversion.jl:600-640)."""
    settings = settings or LossSettings()
    if active == capi.HG_PARAM_N_ZONE:
        ctx.set_active_param(active, zone=zone, n_zones=np.size(params0))
    else:
        ctx.set_active_param(active)
    G = gradient_operator(case.mesh) if settings.slope and active == capi.HG_PARAM_ZB else None
    opt = Adam(lr=lr)
    p = np.array(params0, dtype=np.float64)
    res = InversionResult(params=p)
    for it in range(iters):
        L, parts, g = loss_and_grad(ctx, case, truth, p, active, dt, nsteps, kck, settings, zone=zone, G=G)
        res.loss.append(L)
        res.parts.append(parts)
        res.history.append(p.copy())
        if callback:
            callback(it, L, p)
        p = opt.step(p, g)
    res.params = p
    return res


def ude_settings():
    """The UDE loss: WSE and uv terms only (UDE_bWSE_loss, UDE_b_uv_loss)."""
    return LossSettings(wse=True, uv=True, bound=False, slope=False)


def ude_loss_and_grad(ctx, case, truth, weights, dt, nsteps, kck, settings=None):
    """compute_loss_UDE on the final state after nsteps Euler steps and its
    gradient w.r.t. the network weights (HG_PARAM_UDE must be active).
    truth: WSE_truth, u_truth, v_truth and optionally zb_cell_truth (the bed
    the reference adds to the predicted depth)."""
    settings = settings or ude_settings()
    QN = ctx.step(case.Q0, dt, nsteps, params=weights)
    zb = truth.get("zb_cell_truth", case.zb)
    total, parts, QN_bar, _, _ = final_state_loss(case, QN, truth, settings, zb=zb)
    ctx.adjoint(case.Q0, dt, nsteps, kck, QN_bar, params=weights)
    return total, parts, ctx.lux_grad()


def train_ude(ctx, case, truth, weights0, dt, nsteps, kck=50, lr=0.1, iters=100, settings=None, callback=None):
    """Adam on the reference's UDE loss (the training loop of
    swe_2D_UDE.jl:360-420 with its Adam stage); returns the weights and loss
    history."""
    ctx.set_active_param(capi.HG_PARAM_UDE)
    opt = Adam(lr=lr)
    p = np.array(weights0, dtype=np.float64)
    res = InversionResult(params=p)
    for it in range(iters):
        L, parts, g = ude_loss_and_grad(ctx, case, truth, p, dt, nsteps, kck, settings)
        res.loss.append(L)
        res.parts.append(parts)
        res.history.append(p.copy())
        if callback:
            callback(it, L, p)
        p = opt.step(p, g)
    res.params = p
    return res
