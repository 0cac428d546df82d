/* hgswe.h -- C ABI of libhgswe.so, the MI355X (gfx950) implementation of
 * Hydrograd.jl's 2-D shallow-water right-hand side, explicit Euler step and
 * their reverse-mode adjoint.
 *
 * Conventions
 *  - All indices are 0-based; cells are numbered in the caller's order (the
 *    library may renumber internally for locality and always returns results
 *    in the caller's order).
 *  - State vectors are block SoA of length 3N: [xi(0..N-1), q_x(..), q_y(..)],
 *    exactly like Hydrograd's Q (semi_discretize_swe_2D.jl:93-95).
 *  - Every entry point returns HG_OK (0) or an error code; hg_last_error()
 *    gives the message.  The shim turns a nonzero code into a Julia
 *    error(), mirroring the reference's exceptions.
 *  - A context owns one HIP stream and is not thread-safe: use one context
 *    per caller thread.
 *  - Host-pointer entry points (hg_rhs, hg_step, hg_grad, ...) are
 *    synchronous: they return after results are in host memory.
 */
#ifndef HGSWE_H
#define HGSWE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hg_ctx hg_ctx;

enum {
  HG_OK = 0,
  HG_ERR_ARG = 1,         /* invalid argument / inconsistent array lengths */
  HG_ERR_HIP = 2,         /* HIP runtime error */
  HG_ERR_MESH = 3,        /* malformed mesh (bad CSR, ghost map, clockwise cell) */
  HG_ERR_CONVEYANCE = 4,  /* inlet conveyance <= 1e-10 (bc_2D.jl:678-680) */
  HG_ERR_NONFINITE = 5,   /* NaN / Inf in the state */
  HG_ERR_STATE = 6        /* call order (e.g. no fields set) */
};

/* Boundary kinds (bc_2D.jl:176-239). */
enum { HG_BC_INLETQ = 0, HG_BC_EXITH = 1, HG_BC_WALL = 2, HG_BC_SYMM = 3 };

/* Active parameter of params_vector (semi_discretize_swe_2D.jl:114-199). */
enum {
  HG_PARAM_NONE = 0,   /* params ignored (forward simulation) */
  HG_PARAM_ZB = 1,     /* params = zb per cell; bed data rebuilt (process_bed_2D.jl:46-66) */
  HG_PARAM_N_ZONE = 2, /* params = Manning n per zone; n_cell = params[zone[i]] (process_ManningN_2D.jl:71-98) */
  HG_PARAM_N_CELL = 3, /* params = Manning n per cell */
  HG_PARAM_Q = 4,      /* params = total discharge per inlet-Q boundary */
  HG_PARAM_UDE = 5     /* params = the UDE network's weights (hg_set_lux_ude; the reference
                          passes them as params_vector, semi_discretize_swe_2D.jl:145-158) */
};

/* hg_create flags */
enum {
  HG_REORDER = 1,    /* renumber cells along a Morton curve (needs centroid_xy) */
  HG_CELL_KERNEL = 2 /* one work-item per cell evaluating every face from its side, instead of the
                        default tiled kernel that evaluates intra-tile faces once for both sides
                        (both are bit-identical to the reference's per-side evaluation).  The
                        reverse step follows suit: per cell, every face's VJP from the cell's
                        side, instead of the tiled one that forms each face's VJP for both
                        cells from the side whose normal points to +x (or +y when n_x = 0).
                        The two reverse steps agree to rounding, not bitwise; the tiled one's
                        results do not depend on the tiling when the two sides' normals of a
                        face are exact negatives (as the reference's mesh builder forms them) */
};

/* Mesh topology + geometry, flattened from mesh_2D (mesh_2D.jl:2-71).
 * Cell i owns CSR slots face_off[i] .. face_off[i+1]-1, one per local face
 * in the reference's order (edge node_k -> node_k+1); face_nbr is the
 * neighbour cell or ~g for ghost g; face_nx/ny = cell_normals[i][k];
 * face_len = face_lengths[edge]; area = cell_areas. */
typedef struct {
  int64_t n_cells;
  const int32_t* face_off;    /* [n_cells+1] */
  const int32_t* face_nbr;    /* [face_off[n_cells]] */
  const double* face_nx;      /* [face_off[n_cells]] */
  const double* face_ny;
  const double* face_len;
  const double* area;         /* [n_cells] */
  const double* centroid_xy;  /* [2*n_cells] or NULL; only used by HG_REORDER */
} hg_mesh;

/* Boundary data, flattened from BoundaryConditions2D (bc_2D.jl:3-47).
 * Ghost g sits across CSR slot ghost_slot[g] of its internal cell; its
 * outward normal and length are that slot's. */
typedef struct {
  int64_t n_ghosts;
  const int32_t* ghost_slot;  /* [n_ghosts] */
  const int32_t* ghost_kind;  /* [n_ghosts] HG_BC_* */
  const int32_t* ghost_bc;    /* [n_ghosts] inlet / exit boundary index, -1 for wall / symm */
  int32_t n_inlet;
  int32_t n_exit;
  const int32_t* inlet_off;   /* [n_inlet+1] */
  const int32_t* inlet_ghost; /* ghosts of each inlet-Q boundary in NodeString order */
} hg_bcs;

/* swe_2D_consts (swe_2D_constants.jl:4-18): the scalars the RHS reads. */
typedef struct {
  double g;
  double k_n;
  double h_small;
} hg_consts;

/* ------------------------------------------------------------------ set-up */

/* Build a context: validates and copies the mesh and boundary data to the
 * device (the caller may free its arrays afterwards).  Replaces the set-up
 * of SWE2D_Extra_Parameters (solve_swe_2D.jl:240-277,
 * application_commons.jl:7-44).  With the environment variable HG_GRAPHS=0
 * set when it is created, the context launches every kernel directly
 * instead of replaying cached hipGraphs (same results; for tools that do
 * not follow graph launches). */
int hg_create(hg_ctx** out, const hg_mesh* mesh, const hg_bcs* bcs, const hg_consts* consts, int32_t flags);

int hg_destroy(hg_ctx* ctx);

/* Batched context over n_inst independent meshes (no reference
 * counterpart: the reference runs one case per process).  The instances are
 * laid end to end: instance r's cells, ghosts, inlet-Q and exit-H boundaries
 * follow those of instance r-1, so every other entry point takes
 * concatenated arrays (state [xi; q_x; q_y] over all cells, fields, params,
 * probes) and treats the batch as one block-diagonal mesh advanced with one
 * dt.  Cells are renumbered within each instance. */
int hg_batch_create(hg_ctx** out, int32_t n_inst, const hg_mesh* meshes, const hg_bcs* bcs, const hg_consts* consts,
                    int32_t flags);

/* K parameter fields on one mesh, advanced together (no reference
 * counterpart: the reference runs one parameter field per solve; this is
 * its "many parameter fields" use, e.g. an ensemble of Manning-n fields or
 * of UDE weight sets).  The context behaves exactly like hg_batch_create of
 * n_fields copies of `mesh` -- every other entry point takes and returns the
 * concatenated arrays of the K instances (states [xi; q_x; q_y] over K N
 * cells, per-cell fields and params, probes, inlets and exits of field f
 * after those of field f-1) -- but stores the mesh geometry and its per-cell
 * constants (h_still, zb, S0 and the ghost tables) once, and each step
 * launch covers every field with the work-groups of one tile's K fields on
 * one XCD, so that they read the tile's mesh data from that XCD's L2.  Per
 * field: the state, Manning n, inlet discharges, exit levels and the UDE
 * closure's weight set (hg_set_ude's set_of_cell).  hg_set_fields requires
 * h_still, zb, S0 and the ghost h_still / zb to be the same in every field
 * (HG_ERR_ARG otherwise).  Not available (HG_ERR_STATE): zb, zone n or the
 * network weights as the active parameter, hg_set_n_law, hg_set_lux_ude,
 * hg_jvp, hg_tsit5_solve, hg_batch_set_dt. */
int hg_fields_create(hg_ctx** out, int32_t n_fields, const hg_mesh* mesh, const hg_bcs* bcs, const hg_consts* consts,
                     int32_t flags);

/* Per-instance time steps for a batch (n_inst values, or NULL to return to
 * the scalar dt).  While set, every stepping call (hg_step, hg_advance,
 * hg_adjoint, hg_grad_probes, hg_jvp) advances instance r by dt_inst[r] and
 * ignores its dt argument. */
int hg_batch_set_dt(hg_ctx* ctx, const double* dt_inst);

/* Message of the last error on this context (or of the last failed
 * hg_create when ctx is NULL). */
const char* hg_last_error(const hg_ctx* ctx);

/* Per-cell / per-ghost fields of SWE2D_Extra_Parameters
 * (application_commons.jl:7-44): hstill, hstill_ghostCells, zb_cells,
 * zb_ghostCells, S0_cells (x and y columns), ManningN_cells,
 * inletQ_TotalQ, exitH_WSE. */
int hg_set_fields(hg_ctx* ctx, const double* hstill, const double* hstill_ghost, const double* zb,
                  const double* zb_ghost, const double* S0x, const double* S0y, const double* n,
                  const double* inletQ, const double* exitWSE);

/* Selects how params_vector is read (semi_discretize_swe_2D.jl:114-199).
 * zone_of_cell is required for HG_PARAM_N_ZONE (matID_cells, 0 = default
 * zone; process_SRH_2D_input.jl:136-153). */
int hg_set_active_param(hg_ctx* ctx, int32_t kind, const int32_t* zone_of_cell, int32_t n_zones);

/* ------------------------------------------------------------ forward path */

/* dQdt = swe_2d_rhs(Q, params) (semi_discretize_swe_2D.jl:18-277).
 * Host arrays of length 3N; params has n_params entries per the active
 * parameter (ignored for HG_PARAM_NONE).  Synchronous. */
int hg_rhs(hg_ctx* ctx, const double* Q, const double* params, int64_t n_params, double* dQdt);

/* nsteps calls of custom_ODE_update_cells (custom_ODE_solvers.jl:5-33):
 * Q <- Q + dt * RHS(Q), then xi < h_small => xi = h_small, q = 0.
 * Q is a host array (3N) updated in place; the steps run device-resident
 * with no per-step host traffic (custom_ODE_solve, :36-95, minus its
 * per-step prints and saves). */
int hg_step(hg_ctx* ctx, double* Q, const double* params, int64_t n_params, double dt, int64_t nsteps);

/* Device-resident state: upload / download the 3N state and advance it
 * nsteps with fixed dt (hipGraph-replayed).  Uses the parameters of the
 * last hg_rhs / hg_step / hg_set_params call. */
int hg_set_params(hg_ctx* ctx, const double* params, int64_t n_params);
int hg_state_upload(hg_ctx* ctx, const double* Q);
int hg_state_download(hg_ctx* ctx, double* Q);
int hg_advance(hg_ctx* ctx, double dt, int64_t nsteps);

/* CFL time step: dt = cfl * min over wet cells of (2A/P)/(|u| + sqrt(g h)).
 * No reference counterpart (the reference's CFL field is unused,
 * custom_ODE_solvers.jl:39). */
int hg_cfl_dt(hg_ctx* ctx, const double* Q, double cfl, double* dt);

/* hg_cfl_dt for callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc during a call).  No counterpart in
 * the reference.  Same inputs, output and result bits as hg_cfl_dt; it
 * allocates nothing: the CFL reduction's partial minima and result live in
 * a static device array of 256 KiB per GPU, so meshes of up to 8,388,608
 * cells (larger ones: HG_ERR_ARG).  Calls on one GPU share that array:
 * calls that overlap on one GPU (several host threads, or contexts on
 * several streams) race.  Like hg_cfl_dt, it leaves Q as the device state. */
int hg_cfl_dt_noalloc(hg_ctx* ctx, const double* Q, double cfl, double* dt);

/* Measurement helper with no counterpart in the reference: hg_advance
 * (hipGraph replays) with HIP events recorded on the context stream right
 * before the first and after the last step; *gpu_ms_per_step is the elapsed
 * GPU time over the steps.  When nothing but the step kernel runs per step
 * (no separate conveyance kernel, no closure), it is the step kernel's
 * average launch duration over these launches, back to back. */
int hg_advance_region(hg_ctx* ctx, double dt, int64_t nsteps, double* gpu_ms_per_step);

/* Set-up helper with no counterpart in the reference: instantiates the
 * hipGraph that hg_advance replays for this dt and the dispatch events of
 * hg_advance_timed, without advancing the state, so that a timed call that
 * follows does no graph capture or event creation on the host.  (Changing
 * the launch shape, hg_batch_set_dt or a different dt drops the graph.) */
int hg_prepare_advance(hg_ctx* ctx, double dt);

/* Measurement helper with no counterpart in the reference: hg_advance with
 * every step launched singly and start / stop events attached to each step
 * kernel's dispatch (hipExtLaunchKernelGGL); *kernel_ms is the mean duration
 * of those dispatches, the figure a profiler's kernel trace reports.  The
 * events are created by the first call (for at least 2048 steps) and reused,
 * so later calls of up to that many steps create none; nsteps = 0 only
 * creates them. */
int hg_advance_timed(hg_ctx* ctx, double dt, int64_t nsteps, double* kernel_ms);

/* Timing helper: runs hg_advance(dt, nsteps) `reps` times on the device
 * state and reports the mean time per step measured with HIP events on the
 * context's stream, plus the mean duration of the step kernel alone (200
 * single launches with events on their dispatches). */
int hg_time_advance(hg_ctx* ctx, double dt, int64_t nsteps, int32_t reps, double* ms_per_step,
                    double* ms_step_kernel);

/* Launch shape of the tiled step kernel (no reference counterpart): cells
 * per tile (128, 256 or 512; one work-item per cell) and the work-group to
 * tile mapping: xcd_chunk = 0 maps work-group b to tile b; G > 0 gives each
 * of the 8 XCDs runs of G consecutive tiles within windows of 8G tiles.
 * Every shape gives bit-identical results. */
int hg_set_tile_config(hg_ctx* ctx, int32_t tile_cells, int32_t xcd_chunk);
int hg_get_tile_config(hg_ctx* ctx, int32_t* tile_cells, int32_t* xcd_chunk);

/* Tile size of the tiled reverse step (no reference counterpart): 128, 256
 * or 512 cells, or 0 (default) for the library's choice (256, or 128 when
 * the mesh has fewer than two rounds of three 256-cell work-groups per CU).
 * It shares the step kernel's block -> tile mapping.  Every size gives
 * bit-identical gradients. */
int hg_set_adjoint_tile(hg_ctx* ctx, int32_t tile_cells);

/* Autotuner: after 500 warm-up steps (the device clocks settle), times the
 * step kernel over nsteps steps of every launch shape, twice each (events on
 * each dispatch, after 16 warm-up steps), from the current device state
 * (which is restored afterwards); keeps the shape with the lowest time and
 * reports that mean kernel time in *ms_per_step.  It ends by instantiating
 * the stepping graph for dt with the kept shape and replaying it for 320
 * steps, so the device is warm and no graph is built by the next
 * hg_advance; the state is then restored.  Needs hg_set_fields and a
 * device state (hg_state_upload). */
int hg_autotune(hg_ctx* ctx, double dt, int64_t nsteps, int32_t* tile_cells, int32_t* xcd_chunk,
                double* ms_per_step);

/* ---------------------------------------------------------- adjoint path */

/* Vector-Jacobian product of one swe_2d_rhs evaluation (the pullback Zygote
 * takes through semi_discretize_swe_2D.jl:18-277): given the cotangent of
 * dQdt (3N), returns the cotangents of Q (3N), of the per-cell Manning n
 * (N), of the bed zb (N; zb enters through S0 and zb_ghost rebuilt from zb as
 * on the active-zb path, process_bed_2D.jl:46-66, and through the exit-H
 * depth, bc_2D.jl:763-770) and of the inlet discharges (n_inlet).  Any
 * output pointer may be NULL.  Branch tests and clamps carry no derivative
 * (the semantics ForwardDiff / Zygote give the reference). */
int hg_rhs_vjp(hg_ctx* ctx, const double* Q, const double* params, int64_t n_params, const double* dQdt_bar,
               double* Q_bar, double* n_bar, double* zb_bar, double* inletQ_bar);

/* Reverse-mode adjoint of nsteps Euler steps (custom_ODE_update_cells,
 * custom_ODE_solvers.jl:5-33) from Q0: forward sweep storing a checkpoint
 * every kck steps, then a reverse sweep that recomputes each segment from
 * its checkpoint.  QN_bar is the cotangent of the final state; outputs are
 * the cotangents of Q0, n, zb and the inlet discharges (NULL to skip).
 * kck <= 0 lets the library choose: the whole trajectory kept on the device
 * (nothing recomputed) when it takes at most half of the free device memory,
 * else ceil(sqrt(nsteps)).  Results are bitwise the same for every kck. */
int hg_adjoint(hg_ctx* ctx, const double* Q0, const double* params, int64_t n_params, double dt, int64_t nsteps,
               int64_t kck, const double* QN_bar, double* Q0_bar, double* n_bar, double* zb_bar, double* inletQ_bar);

/* hg_adjoint in two calls, for a loss formed by the caller from the final
 * state (swe_2D_inversion.jl:247-483 and swe_2D_UDE.jl form theirs on the
 * host): hg_adjoint_forward runs and records the forward sweep of hg_adjoint
 * (QN: the final state, caller order; NULL to skip), hg_adjoint_reverse then
 * takes the final state's cotangent and returns what hg_adjoint returns,
 * without running the forward sweep again.  The reverse must follow its
 * forward with no other gradient call and no call that sets fields or
 * parameters in between (else HG_ERR_STATE); one reverse per forward.  The
 * results are bitwise those of hg_adjoint. */
int hg_adjoint_forward(hg_ctx* ctx, const double* Q0, const double* params, int64_t n_params, double dt,
                       int64_t nsteps, int64_t kck, double* QN);
int hg_adjoint_reverse(hg_ctx* ctx, const double* QN_bar, double* Q0_bar, double* n_bar, double* zb_bar,
                       double* inletQ_bar);

/* The checkpoint interval the last hg_adjoint / hg_grad_probes call used
 * (what kck <= 0 resolved to; 0 before any such call).  No reference
 * counterpart. */
int hg_last_checkpoint_interval(const hg_ctx* ctx, int64_t* kck);

/* Gradient of the probe loss L = sum_p (xi_p + h_still_p + zb_p - target_p)^2
 * at the final state (the WSE misfit of swe_2D_inversion.jl:387-399 without
 * its normalisation) after nsteps steps from Q0, with respect to n, zb and
 * the inlet discharges; one checkpointed forward + reverse sweep.  As in
 * hg_adjoint and hg_rhs_vjp, a NULL n_bar or zb_bar skips the accumulation of
 * that cotangent (the reverse step then reads and writes less); kck as for
 * hg_adjoint. */
int hg_grad_probes(hg_ctx* ctx, const double* Q0, const double* params, int64_t n_params, double dt, int64_t nsteps,
                   int64_t kck, int64_t n_probe, const int32_t* probe, const double* target, double* loss,
                   double* n_bar, double* zb_bar, double* inletQ_bar);

/* Adaptive Tsit5 (Tsitouras 5(4) pair, first-same-as-last) over the device
 * RHS: solve(ODEProblem(ode_f, Q0, (t0, t1), p), Tsit5(); abstol, reltol,
 * saveat) as the reference's examples run it (swe_2D_forward_simulation.jl:
 * 24-53).  Saves n_save + 1 states at t0 + k (t1 - t0) / n_save (the steps
 * land on the save times) into Qsave ((n_save + 1) x 3N, caller order).
 * Error norm: RMS of err_i / (abstol + reltol max(|y_i|, |y_new_i|)), as
 * OrdinaryDiffEq's default.  dt0 is the first trial step.  stats (optional):
 * accepted steps, rejected steps, RHS evaluations.  No dry clamp is applied
 * (that belongs to the custom Euler driver). */
int hg_tsit5_solve(hg_ctx* ctx, const double* Q0, const double* params, int64_t n_params, double t0, double t1,
                   int64_t n_save, double abstol, double reltol, double dt0, double* Qsave, int64_t* stats);

/* The Tsit5 coefficients the solver uses: c[7], a[7][6] (row i = stage i+1),
 * btilde[7] (fifth-order minus embedded fourth-order weights; the fifth-order
 * weights are row 6 of a).  Host-only, no device needed. */
int hg_tsit5_tableau(double* c, double* a, double* btilde);

/* Forward mode (JVP) of nsteps Euler steps (custom_ODE_update_cells) along
 * one tangent direction of the initial state (dQ0, 3N, NULL = 0) and of the
 * active parameter (dparams, n_params, NULL = 0): QN and its tangent dQN.
 * The device counterpart of ForwardDiff.Dual through the step
 * (swe_2D_sensitivity.jl:33-80, ForwardDiff.jacobian column by column);
 * branch tests follow the primal values, clamps have zero tangent.  With
 * the reference's UDE network on (hg_set_lux_ude), the network runs on dual
 * numbers every step: its output's tangent follows the state tangent and,
 * with HG_PARAM_UDE active, the weight tangent dparams (the reference's
 * AutoForwardDiff / ForwardDiffSensitivity UDE training, swe_2D_UDE.jl:
 * 156-157).  The MLP closure of hg_set_ude has no forward mode
 * (HG_ERR_STATE). */
int hg_jvp(hg_ctx* ctx, const double* Q0, const double* dQ0, const double* params, int64_t n_params,
           const double* dparams, double dt, int64_t nsteps, double* QN, double* dQN);

/* ------------------------------------------------------- UDE closure */

/* Manning's n from a neural closure, evaluated on the matrix cores every
 * RHS evaluation: n_i = 0.01 + softplus(MLP(h_i, |u|_i, zb_i)), MLP =
 * 3 -> 16 -> 16 -> 1 with tanh hidden layers.  It has the role of the
 * reference's UDE ManningN (update_ManningN_UDE, process_ManningN_2D.jl:
 * 216-273, called from semi_discretize_swe_2D.jl:165-180) with this network.
 * Features: h = xi + h_still clamped to h_small (then q = 0), |u| =
 * sqrt(u^2 + v^2 + eps) with u = q_x / h, v = q_y / h, and zb.
 * weights: n_sets x 353 values, per set W1[16][3], b1[16], W2[16][16]
 * (row = output unit), b2[16], W3[16], b3.  set_of_cell (caller order, NULL =
 * set 0 everywhere) picks each cell's set, e.g. one per batched instance.
 * precision: 32 (float, the production setting) or 64 (double).  At most
 * 65535 sets.  n_sets = 0 turns the closure off and restores the n of
 * hg_set_fields.  While on, n cannot be the active parameter. */
int hg_set_ude(hg_ctx* ctx, int32_t n_sets, const double* weights, const int32_t* set_of_cell, int32_t precision);

/* The closure's n (N, caller order) for state Q (3N). */
int hg_ude_n(hg_ctx* ctx, const double* Q, double* n);

/* Cotangents of the weights (n_sets x 353, set-major) accumulated by the
 * last hg_rhs_vjp, hg_adjoint or hg_grad_probes call with the closure on.
 * Each set's cotangent collects the cells that use it, so one reverse sweep
 * over a batch with one weight set per instance gives every set's gradient
 * of the summed loss. */
int hg_ude_grad(hg_ctx* ctx, double* w_bar);

/* ---- Variable Manning n: the forward simulation's "variable" option ---- */

/* Law kinds of create_manning_function (process_ManningN_2D.jl:122-134). */
enum {
  HG_NLAW_OFF = 0,
  HG_NLAW_POWER = 1,     /* n_lower + (n_upper - n_lower) (h + eps)^-k              (:154-165) */
  HG_NLAW_SIGMOID = 2,   /* n_lower + (n_upper - n_lower) / (1 + exp(k (h - h_mid))) (:170-182) */
  HG_NLAW_INVERSE = 3,   /* n_lower + (n_upper - n_lower) / (1 + k h)                (:139-150) */
  HG_NLAW_H_UMAG_KS = 4  /* Cheng (2008) friction factor of h, |U|, ks -> n          (:186-213) */
};

/* Replaces update_ManningN_forward_simulation inside swe_2d_rhs
 * (semi_discretize_swe_2D.jl:129-135, process_ManningN_2D.jl:102-117): from
 * now on every RHS evaluation (hg_rhs, hg_step, hg_advance, hg_tsit5_solve,
 * and the adjoint / JVP) first sets n per cell from the state by the law,
 * with h = xi + h_still clamped to h_small and |U| = |q| / h, and the inlet
 * conveyance and friction use that n.
 * params: {n_lower, n_upper, k, h_mid} (h_mid is read by the sigmoid only;
 * all four are ignored by HG_NLAW_H_UMAG_KS).  ks: roughness height per cell
 * (N, caller order; ks_cells of setup_ManningN, process_ManningN_2D.jl:56-61),
 * required by HG_NLAW_H_UMAG_KS and ignored otherwise.  Errors like the
 * reference's assertions: k > 0 (HG_ERR_ARG), h_mid > 0 for the sigmoid.
 * kind = HG_NLAW_OFF restores the n of hg_set_fields.  The law excludes the
 * UDE closure and Manning n as the active parameter (HG_ERR_STATE). */
int hg_set_n_law(hg_ctx* ctx, int32_t kind, const double* params, const double* ks);

/* The law at given depths and velocity magnitudes (N each, caller order):
 * n and the reference's auxiliary outputs h/ks, friction factor and Reynolds
 * number (update_ManningN_forward_simulation's four results,
 * process_ManningN_2D.jl:102-117; zero for the depth-only laws).  Any output
 * may be NULL.  Evaluated on the device by the same code as the RHS. */
int hg_n_law_eval(hg_ctx* ctx, const double* h, const double* Umag, double* n, double* h_ks, double* f,
                  double* Re);

/* Manning n per cell (N, caller order) that the RHS uses at state Q (3N):
 * the law's or the UDE closure's when one is on, else the current field. */
int hg_cell_n(hg_ctx* ctx, const double* Q, double* n);

/* ---- The reference's UDE network (UDE_choice, Lux model of create_NN_model) ---- */

enum {
  HG_LUX_OFF = 0,
  HG_LUX_MANNING_H = 1,         /* n = NN(h)                 (process_ManningN_2D.jl:216-273) */
  HG_LUX_MANNING_H_UMAG_KS = 2, /* n = NN(h, |U|, ks)        (same) */
  HG_LUX_FLOW_RESISTANCE = 3    /* friction magnitude = NN(h, |q|) (process_FlowResistance_UDE.jl:3-46,
                                   friction m q / sqrt(q_x^2 + q_y^2 + eps), semi_discretize_swe_2D.jl:521-531) */
};
/* Hidden-layer activations (get_activation, process_UDE.jl:93-107; NNlib definitions). */
enum { HG_ACT_RELU = 0, HG_ACT_LEAKYRELU = 1, HG_ACT_SIGMOID = 2, HG_ACT_TANH = 3, HG_ACT_SOFTPLUS = 4 };

/* The network of create_NN_model (process_UDE.jl:2-65): per hidden layer
 * Dense(in, width[l], act[l]) -> LayerNorm(width[l]); then Dense(width[last], 1)
 * and lo + (hi - lo) sigmoid(.) with output_bounds = {lo, hi}.  Inputs are
 * normalised to [-1, 1] by their bounds: h by h_bounds; |U| by u_bounds
 * (Umag_bounds, ManningN_h_Umag_ks); |q| by {h_lo u_lo, h_hi u_hi}
 * (velocity_magnitude_bounds, FlowResistance); ks by ks_bounds.
 * ln_epsilon: LayerNorm epsilon, 0 for Lux's default Float32(1e-5). */
typedef struct {
  int32_t choice;          /* HG_LUX_* */
  int32_t n_hidden;        /* 1..4 hidden layers */
  int32_t width[4];        /* 1..16 units each */
  int32_t act[4];          /* HG_ACT_* each */
  double output_bounds[2];
  double h_bounds[2];
  double u_bounds[2];
  double ks_bounds[2];
  double ln_epsilon;
} hg_lux_net;

/* Number of weights of the network (or -1 if the description is invalid).
 * Flat order, per hidden layer l: Dense weight (width[l] x in, column-major,
 * as Lux stores it), Dense bias, LayerNorm scale, LayerNorm bias; then the
 * output Dense weight (width[last]) and bias.  The Julia binding builds this
 * vector from Lux's parameter NamedTuple by field name. */
int64_t hg_lux_n_params(const hg_lux_net* net);

/* Turns the network on (net NULL: off, restoring the n of hg_set_fields).
 * From then on every RHS evaluation first evaluates the network per cell
 * from the state (clamped h, masked q) and uses its output as Manning n
 * (inlet conveyance and friction) or, for HG_LUX_FLOW_RESISTANCE, as the
 * friction magnitude (the inlet conveyance keeps the field n).  params: the
 * weights (hg_lux_n_params of them); ks: per cell (caller order), required by
 * HG_LUX_MANNING_H_UMAG_KS.  Excludes the MLP closure, the variable-n laws
 * and n as the active parameter.  With HG_PARAM_UDE active, the params of
 * each call replace the weights. */
int hg_set_lux_ude(hg_ctx* ctx, const hg_lux_net* net, const double* params, const double* ks);

/* The network's output per cell (N, caller order) at state Q (3N). */
int hg_lux_eval(hg_ctx* ctx, const double* Q, double* out);

/* Cotangent of the weights accumulated by the last hg_rhs_vjp, hg_adjoint or
 * hg_grad_probes call with the network on. */
int hg_lux_grad(hg_ctx* ctx, double* params_bar);

/* Self-check with no counterpart in the reference: D = A B on the matrix
 * cores for A (16x4) and B (4x16) row-major, D (16x16) row-major, in the
 * given precision (32 or 64); verifies the lane layout the closure uses. */
int hg_mfma_probe(int32_t precision, const double* A, const double* B, double* D);

/* Measurement helper with no counterpart in the reference: copies nbytes
 * between two scratch device buffers `reps` times with a plain streaming
 * kernel and reports the achieved bandwidth (read + write bytes / time).  Its
 * known traffic calibrates rocprof's FETCH_SIZE / WRITE_SIZE counters. */
int hg_probe_copy(hg_ctx* ctx, int64_t nbytes, int32_t reps, double* gbps);

/* Library / device information string (arch, CU count, build flags). */
const char* hg_info(void);

#ifdef __cplusplus
}
#endif

#endif /* HGSWE_H */
