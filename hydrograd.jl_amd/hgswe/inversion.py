"""Parameter inversion and UDE training over the device path: the
reference's inversion loss on the final state (compute_loss_inversion,
swe_2D_inversion.jl:247-483) and its UDE loss (compute_loss_UDE,
swe_2D_UDE.jl:522-599: the same WSE and uv terms), their gradients through
the library's adjoint (hg_adjoint; hg_lux_grad for the network weights), and
the optimiser loop the reference drives through Optimization.jl
(swe_2D_inversion.jl:600-732, swe_2D_UDE.jl:263-320): Adam and Descent
(Optimisers.jl) and LBFGS with a backtracking line search (Optim.jl,
LineSearches.jl), run as stages in the reference's order.

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
class Descent:
    """Optimisers.jl Descent: p -= lr g (the reference's "Descent" choice,
    swe_2D_inversion.jl:683-684)."""
    lr: float = 0.1

    def step(self, p, g):
        return p - self.lr * g


class LineSearchError(RuntimeError):
    pass


def backtracking(phi, phi0, dphi0, alpha=1.0, c1=1e-4, rho_hi=0.5, rho_lo=0.1, max_iter=1000):
    """Backtracking line search on phi(a) = f(x + a d) for the Armijo
    condition phi(a) <= phi0 + c1 a dphi0 (dphi0 = g.d < 0); the reference's
    LBFGS(linesearch = BackTracking()) (swe_2D_inversion.jl:687,
    LineSearches.jl, third order).  The first reduction minimises the
    quadratic through phi0, dphi0 and phi(a); later ones the cubic through
    phi0, dphi0 and the last two trial values (Nocedal & Wright, Numerical
    Optimization, 2nd ed., Sec. 3.5); each new trial is kept in
    [rho_lo a, rho_hi a].  A non-finite trial value counts as a failed
    Armijo test.  phi returns (value, payload); returns (a, value,
    payload)."""
    a_prev, a = alpha, alpha
    f_prev = phi0
    f_a, pay = phi(a)
    it = 0
    while not np.isfinite(f_a) or f_a > phi0 + c1 * a * dphi0:
        it += 1
        if it > max_iter or a < np.sqrt(EPS64):
            raise LineSearchError(f"backtracking: no sufficient decrease after {it - 1} reductions (a = {a:g})")
        if it == 1 or not np.isfinite(f_a) or not np.isfinite(f_prev):
            denom = 2.0 * (f_a - phi0 - dphi0 * a)
            a_new = -(dphi0 * a * a) / denom if np.isfinite(denom) and denom > 0 else rho_hi * a
        else:
            r1 = f_a - phi0 - dphi0 * a
            r0 = f_prev - phi0 - dphi0 * a_prev
            div = 1.0 / (a_prev * a_prev * a * a * (a - a_prev))
            ca = (a_prev * a_prev * r1 - a * a * r0) * div
            cb = (-a_prev ** 3 * r1 + a ** 3 * r0) * div
            if ca == 0.0:
                a_new = -dphi0 / (2.0 * cb)
            else:
                disc = max(cb * cb - 3.0 * ca * dphi0, 0.0)
                a_new = (-cb + np.sqrt(disc)) / (3.0 * ca)
        if not np.isfinite(a_new):
            a_new = rho_hi * a
        a_prev, f_prev = a, f_a
        a = max(min(a_new, rho_hi * a), rho_lo * a)
        f_a, pay = phi(a)
    return a, f_a, pay


@dataclass
class LBFGS:
    """Limited-memory BFGS with the backtracking line search above (the
    reference's LBFGS(linesearch = BackTracking()), Optim.jl defaults:
    m = 10 pairs, unit first trial step, initial inverse Hessian scaled by
    s.y / y.y of the newest pair).  The direction is the two-loop recursion
    (Nocedal & Wright, Alg. 7.4).  A pair with s.y <= 0 is not stored, and a
    direction that is not one of descent restarts from -g with the history
    cleared.  Stops when max |g| <= g_tol (retcode "Success"), when the
    relative loss change falls below f_reltol (> 0; "Success"), when the
    line search finds no decrease ("Failure", at the last accepted point) or
    after maxiters iterations ("MaxIters")."""
    m: int = 10
    g_tol: float = 1e-8
    f_reltol: float = 0.0
    alpha0: float = 1.0

    def direction(self, g, S, Y):
        q = g.copy()
        al = []
        for s, y in zip(reversed(S), reversed(Y)):
            a = (s @ q) / (y @ s)
            al.append(a)
            q -= a * y
        if S:
            q *= (S[-1] @ Y[-1]) / (Y[-1] @ Y[-1])
        for (s, y), a in zip(zip(S, Y), reversed(al)):
            b = (y @ q) / (y @ s)
            q += s * (a - b)
        return -q

    def minimize(self, fg, x0, maxiters, callback=None):
        """fg(x) -> (f, g).  Returns (x, f, retcode, iterations)."""
        x = np.array(x0, dtype=np.float64)
        f, g = fg(x)
        S, Y = [], []
        for it in range(maxiters):
            if np.max(np.abs(g)) <= self.g_tol:
                return x, f, "Success", it
            d = self.direction(g, S, Y)
            dphi0 = float(g @ d)
            if not dphi0 < 0.0:
                S.clear()
                Y.clear()
                d = -g
                dphi0 = -float(g @ g)

            def phi(a, x=x, d=d):
                fa, ga = fg(x + a * d)
                return fa, ga

            try:
                a, f_new, g_new = backtracking(phi, f, dphi0, alpha=self.alpha0)
            except LineSearchError:  # no decrease measurable along d: stop where we are
                return x, f, "Failure", it
            s = a * d
            y = g_new - g
            if s @ y > 0.0:
                S.append(s)
                Y.append(y)
                if len(S) > self.m:
                    S.pop(0)
                    Y.pop(0)
            x = x + s
            f_old, f, g = f, f_new, g_new
            if callback:
                callback(it, f, x)
            if self.f_reltol > 0.0 and abs(f - f_old) <= self.f_reltol * abs(f):
                return x, f, "Success", it + 1
        return x, f, "MaxIters", maxiters


def _stage_fg(fg):
    """Wraps a loss-and-gradient function for the line search: a trial
    point the solver rejects (an inlet conveyance at or below 1e-10, or a
    non-finite state: the reference's assertion and divergence) evaluates to
    +inf, so the search backs off from it."""
    def wrapped(x):
        try:
            return fg(x)
        except capi.HGError as e:
            if e.code in (capi.HG_ERR_CONVEYANCE, capi.HG_ERR_NONFINITE):
                return np.inf, np.full(np.size(x), np.nan)
            raise
    return wrapped


def run_optimizers(fg, p0, stages, callback=None):
    """The reference's optimiser loop (swe_2D_inversion.jl:664-729,
    swe_2D_UDE.jl:263-320): stages = [(choice, learning rate, max
    iterations), ...] with choice "Adam", "Descent" or "LBFGS" (the learning
    rate is ignored by LBFGS).  Each stage starts from the previous one's
    parameters; a stage with max iterations <= 0 is skipped; a stage that
    converges (LBFGS "Success") ends the loop.  fg(p) -> (loss, grad);
    callback(stage, iteration, loss, p).  Returns (p, [(choice, retcode,
    iterations), ...])."""
    p = np.array(p0, dtype=np.float64)
    log = []
    for k, (choice, lr, iters) in enumerate(stages):
        if iters <= 0:
            continue
        cb = (lambda it, L, x, k=k: callback(k, it, L, x)) if callback else None
        if choice in ("Adam", "Descent"):
            opt = Adam(lr=lr) if choice == "Adam" else Descent(lr=lr)
            for it in range(iters):
                L, g = fg(p)
                if cb:
                    cb(it, L, p)
                p = opt.step(p, g)
            log.append((choice, "MaxIters", iters))
        elif choice == "LBFGS":
            p, _, rc, n = LBFGS().minimize(_stage_fg(fg), p, iters, callback=cb)
            log.append((choice, rc, n))
            if rc == "Success":
                break
        else:
            raise ValueError(f"invalid optimizer choice {choice!r}: Adam, Descent or LBFGS")
    return p, log


@dataclass
class InversionResult:
    params: np.ndarray
    loss: list = field(default_factory=list)
    parts: list = field(default_factory=list)
    history: list = field(default_factory=list)


def loss_and_grad(ctx, case, truth, params, active, dt, nsteps, kck, settings: LossSettings, zone=None, G=None):
    """Loss of the final state after nsteps Euler steps and its gradient
    w.r.t. the active parameter vector (zone n, per-cell n, zb or inlet Q).
    One recorded forward sweep serves both the loss and the reversal."""
    QN = ctx.adjoint_forward(case.Q0, dt, nsteps, kck, params=params)
    zb = params if active == capi.HG_PARAM_ZB else case.zb
    total, parts, QN_bar, pb, zbd = final_state_loss(case, QN, truth, settings, params=params, zb=zb, G=G)
    wrt = {capi.HG_PARAM_N_ZONE: ("n",), capi.HG_PARAM_N_CELL: ("n",), capi.HG_PARAM_ZB: ("zb",),
           capi.HG_PARAM_Q: ("Q",)}.get(active, ())
    _, nb, zbb, qb = ctx.adjoint_reverse(QN_bar, wrt=wrt)
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


def invert(ctx, case, truth, params0, active, dt, nsteps, kck=0, lr=1e-3, iters=100, settings=None, zone=None,
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


def invert_staged(ctx, case, truth, params0, active, dt, nsteps, stages, kck=0, settings=None, zone=None):
    """The reference's optimiser schedule on the inversion loss, e.g.
    stages = [("Adam", 2e-3, 60), ("LBFGS", None, 20)] as the examples'
    run_control.json lists ["Adam", "LBFGS"] (swe_2D_inversion.jl:664-729).
    Every loss evaluation (line-search trials included) is recorded."""
    settings = settings or LossSettings()
    if active == capi.HG_PARAM_N_ZONE:
        ctx.set_active_param(active, zone=zone, n_zones=np.size(params0))
    else:
        ctx.set_active_param(active)
    G = gradient_operator(case.mesh) if settings.slope and active == capi.HG_PARAM_ZB else None
    res = InversionResult(params=np.array(params0, dtype=np.float64))
    stage_loss = []

    def fg(p):
        L, parts, g = loss_and_grad(ctx, case, truth, p, active, dt, nsteps, kck, settings, zone=zone, G=G)
        res.loss.append(L)
        res.parts.append(parts)
        res.history.append(np.array(p, dtype=np.float64))
        return L, g

    def cb(stage, it, L, p):
        stage_loss.append((stage, it, L))

    res.params, res.stages = run_optimizers(fg, params0, stages, callback=cb)
    res.stage_loss = stage_loss
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
    QN = ctx.adjoint_forward(case.Q0, dt, nsteps, kck, params=weights)
    zb = truth.get("zb_cell_truth", case.zb)
    total, parts, QN_bar, _, _ = final_state_loss(case, QN, truth, settings, zb=zb)
    ctx.adjoint_reverse(QN_bar, wrt=())
    return total, parts, ctx.lux_grad()


def train_ude(ctx, case, truth, weights0, dt, nsteps, kck=0, lr=0.1, iters=100, settings=None, callback=None):
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


def train_ude_staged(ctx, case, truth, weights0, dt, nsteps, stages, kck=0, settings=None):
    """The reference's UDE optimiser schedule (swe_2D_UDE.jl:263-320), e.g.
    stages = [("Adam", 0.1, 100), ("LBFGS", None, 50)]; every loss
    evaluation is recorded."""
    ctx.set_active_param(capi.HG_PARAM_UDE)
    res = InversionResult(params=np.array(weights0, dtype=np.float64))
    stage_loss = []

    def fg(p):
        L, parts, g = ude_loss_and_grad(ctx, case, truth, p, dt, nsteps, kck, settings)
        res.loss.append(L)
        res.parts.append(parts)
        res.history.append(np.array(p, dtype=np.float64))
        return L, g

    res.params, res.stages = run_optimizers(fg, weights0, stages,
                                            callback=lambda k, it, L, p: stage_loss.append((k, it, L)))
    res.stage_loss = stage_loss
    return res
