#!/usr/bin/env python3
"""Benchmark: forward Euler steps of the 2-D shallow-water solver on the
~1M-triangle channel (10 km x 1 km, seeded Delaunay mesh, bumpy bed, 5 Manning
zones, INLET-Q 500 m^3/s, EXIT-H 3 m, walls), FP64, dt from CFL 0.5.

A step is one fused RHS + Euler update of every cell (custom_ODE_update_cells);
the state and all mesh data are resident in HBM before the timed region.
Prints one JSON line (see the keys below).  Single GPU only.
"""
import argparse
import json

import numpy as np
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hydrograd.jl_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def bytes_per_cell_step(mesh, fields=1):
    """Algorithmic HBM bytes of one forward cell-step: read xi, q_x, q_y (24),
    write them (24), per-cell constants h_still, zb, n, S0x, S0y, area (48),
    per cell-face neighbour index (4) + n_x, n_y (16) + length (8).  With K
    fields sharing the mesh, each field reads its own state and n (56 B) and
    the mesh data (h_still, zb, S0x, S0y, area and the faces) is read once
    for all K: 40 + 28 x degree over K."""
    avg_deg = float(mesh.face_off[-1]) / mesh.n_cells
    if fields == 1:
        return 24 + 24 + 48 + 28 * avg_deg
    return 24 + 24 + 8 + (40 + 28 * avg_deg) / fields


def bytes_per_gradient_cell_step(mesh, kck, steps=None):
    """Algorithmic HBM bytes of one horizon cell-step of a gradient w.r.t.
    n: the forward sweep (one forward cell-step), the reverse step (read
    Q_n 24, read lambda_{n+1} 24, write lambda_n 24, per-cell constants 48,
    faces 28 per cell-face, read and write the n cotangent 16) and either
    the segment recompute (one more forward cell-step) plus the checkpoint
    write (24 / kck), or, when one segment covers the run (kck >= steps),
    the store of every state (24)."""
    avg_deg = float(mesh.face_off[-1]) / mesh.n_cells
    reverse = 24 + 24 + 24 + 48 + 28 * avg_deg + 16
    if steps is not None and kck >= steps:
        return bytes_per_cell_step(mesh) + 24 + reverse
    return 2 * bytes_per_cell_step(mesh) + reverse + 24.0 / kck


def gradient_roofline(mesh, cells, steps, kck, seconds):
    bpc = bytes_per_gradient_cell_step(mesh, kck, steps)
    achieved = bpc * cells * steps / seconds / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "bytes_per_horizon_cell_step": bpc}


def traffic_from_profiles():
    """HBM bytes per launch of the step kernel from the committed rocprof
    counter summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel
    calibration), or None when not measured."""
    p = os.path.join(ROOT, "profiles", "k_swe_tile_channel1M_traffic.txt")
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def host_threads():
    """Host threads this process may use: its CPU affinity, capped by
    OMP_NUM_THREADS when that is set."""
    n = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    return min(n, int(omp)) if omp.isdigit() and int(omp) > 0 else n


def timed_sample(run, seconds, cap):
    """Times run(k) for a k sized to about `seconds` from a one-step probe."""
    t0 = time.perf_counter()
    run(1)
    per = time.perf_counter() - t0
    n = max(1, min(cap, int(seconds / max(per, 1e-3))))
    t0 = time.perf_counter()
    run(n)
    return n, time.perf_counter() - t0


def cpu_baseline(case, seconds):
    """The oracle (C++ restatement of the reference's loop structure) on the
    host, one thread like the single-threaded reference, on a bounded sample
    of the same workload: a few Euler steps of the same mesh.  The same
    sample on all host threads (OpenMP over cells) is reported beside it."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    orc = pyoracle.Oracle(case)
    pyoracle.set_threads(1)
    n, el = timed_sample(lambda k: orc.run(case.Q0, case.dt, k), seconds, 50)
    out = {"value": case.N * n / el, "unit": "cell-steps/s", "cores": 1, "kind": "port",
           "sample": f"{n} Euler steps of the same {case.N}-cell mesh, C++ oracle -O3, 1 thread"}
    nt = host_threads()
    pyoracle.set_threads(nt)
    n, el = timed_sample(lambda k: orc.run(case.Q0, case.dt, k), seconds / 4, 50)
    pyoracle.set_threads(1)
    out["all_cores"] = {"value": case.N * n / el, "cores": nt, "sample": f"{n} Euler steps, {nt} threads"}
    return out


def gradient_case(cells, steps, seed=2, n_probe=200):
    """Adjoint workload: the channel generator at ~`cells` triangles, per-cell
    n = 0.03 as the initial guess, probe targets from a forward run with a
    'true' lognormal n (mean 0.035, sigma 0.3), 200 seeded probe cells."""
    from hgswe import cases

    c = cases.channel(cells, seed=seed, nsteps=steps)
    rng = np.random.default_rng(seed + 100)
    sig = 0.3
    n_true = np.exp(rng.normal(np.log(0.035) - 0.5 * sig * sig, sig, c.N))
    probe = rng.choice(c.N, size=n_probe, replace=False).astype(np.int32)
    return c.copy_fields(n=np.full(c.N, 0.03)), n_true, probe


def cpu_gradient_baseline(case, probe, target, seconds):
    """CPU cost of a gradient: the oracle's forward Euler steps plus a
    forward-mode (Dual) tangent sweep of the same steps, one thread.  A CPU
    reverse-mode gradient needs at least one forward and one sweep of
    tangent-linear cost, so this is a lower bound on its time; the oracle's
    own reverse mode (a tape, as an operator-overloading AD tool builds) is
    timed beside it as `reverse_tape`."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    orc = pyoracle.Oracle(case)
    dn = np.ones(case.N) * 1e-3

    def run(k):
        orc.run(case.Q0, case.dt, k)
        orc.jvp_run(case.Q0, case.dt, k, dn=dn, probe=probe, target=target)

    pyoracle.set_threads(1)
    n, el = timed_sample(run, seconds, 200)
    out = {"value": case.N * n / el, "unit": "cell-steps/s", "cores": 1, "kind": "port",
           "sample": f"{n} steps of the {case.N}-cell gradient mesh: C++ oracle forward + Dual tangent sweep "
                     "(lower bound on a CPU reverse-mode gradient), 1 thread"}
    nt = host_threads()
    pyoracle.set_threads(nt)
    n, el = timed_sample(run, seconds / 4, 200)
    pyoracle.set_threads(1)
    out["all_cores"] = {"value": case.N * n / el, "cores": nt, "sample": f"{n} steps, {nt} threads"}
    # a measured CPU reverse-mode gradient of the same loss: the oracle's tape
    # (operator-overloading) reverse mode -- the forward run keeping every
    # state, then each step recorded from its stored input and swept
    # backwards -- over a bounded number of steps, 1 thread
    try:  # an auxiliary figure: a failure is reported in the line, not fatal to it
        n, el = timed_sample(lambda k: orc.grad_run(case.Q0, case.dt, k, probe, target, wrt=("n",)), seconds / 2, 16)
        out["reverse_tape"] = {"value": case.N * n / el, "cores": 1,
                               "sample": f"{n} steps of the {case.N}-cell gradient mesh: C++ oracle tape reverse "
                                         "mode (forward run + per-step recording and reverse sweep), 1 thread"}
    except Exception as e:
        out["reverse_tape"] = {"error": repr(e)}
    finally:
        orc.release_tape()
    return out


def measure_gradient(args, capi):
    c, n_true, probe = gradient_case(args.grad_cells, args.grad_steps)
    ctx = capi.Context(c)
    ctx.set_active_param(capi.HG_PARAM_N_CELL)
    if not (args.no_tune or args.tile_config):
        ctx.set_params(n_true)
        ctx.upload(c.Q0)
        T, x, _ = ctx.autotune(c.dt, 200)
        print(f"# gradient autotune: {T}-cell tiles, xcd_chunk={x}", file=sys.stderr, flush=True)
    QT = ctx.step(c.Q0, c.dt, args.grad_steps, params=n_true)
    target = QT[probe] + c.hstill[probe] + c.zb[probe]
    ctx.grad_probes(c.Q0, c.dt, 20, args.grad_kck, probe, target, params=c.n, wrt=("n",))  # warm-up
    t0 = time.perf_counter()
    loss, gn, _, _ = ctx.grad_probes(c.Q0, c.dt, args.grad_steps, args.grad_kck, probe, target, params=c.n,
                                     wrt=("n",))
    el = time.perf_counter() - t0
    out = {"value": c.N * args.grad_steps / el, "unit": "cell-steps/s", "ms_per_gradient": el * 1e3,
           "config": {"workload": "channel_100k_tri_probe_loss_gradient", "cells": c.N, "steps": args.grad_steps,
                      "checkpoint_every": args.grad_kck, "probes": int(probe.size), "wrt": "per-cell Manning n",
                      "dt": c.dt},
           "roofline": gradient_roofline(c.mesh, c.N, args.grad_steps, args.grad_kck, el),
           "loss": loss, "grad_norm": float(np.linalg.norm(gn))}
    out["stored_trajectory"] = stored_gradient(ctx, c, args.grad_steps, probe, target)
    if not args.no_cpu:
        cb = cpu_gradient_baseline(c, probe, target, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = out["value"] / cb["value"]
        out["speedup_vs_cpu_all_cores"] = out["value"] / cb["all_cores"]["value"]
        if "value" in cb["reverse_tape"]:
            out["speedup_vs_cpu_reverse_tape"] = out["value"] / cb["reverse_tape"]["value"]
    return out


def measure_gradient_full(args, capi, case, tile):
    """The probe-loss gradient w.r.t. per-cell n on the forward workload's
    own 1M-triangle mesh: 1,000 steps (the forward workload's length,
    whatever --steps says), 200 probes, targets from a forward run with a
    'true' lognormal n (mean 0.035, sigma 0.3), initial guess n = 0.03.  The
    checkpoint interval is --kck (default 0, the library's choice: the whole
    trajectory when it fits in half the free HBM, which 24 GB does), and the
    same gradient is also timed with a checkpoint every --grad-kck steps."""
    rng = np.random.default_rng(102)
    sig = 0.3
    n_true = np.exp(rng.normal(np.log(0.035) - 0.5 * sig * sig, sig, case.N))
    probe = rng.choice(case.N, size=200, replace=False).astype(np.int32)
    c = case.copy_fields(n=np.full(case.N, 0.03))
    ctx = capi.Context(c)
    ctx.set_active_param(capi.HG_PARAM_N_CELL)
    ctx.set_tile_config(*tile)
    steps = args.grad_1m_steps
    QT = ctx.step(c.Q0, c.dt, steps, params=n_true)
    target = QT[probe] + c.hstill[probe] + c.zb[probe]

    def timed(kck):
        ctx.grad_probes(c.Q0, c.dt, steps, kck, probe, target, params=c.n, wrt=("n",))  # warm-up (allocates)
        t0 = time.perf_counter()
        loss, gn, _, _ = ctx.grad_probes(c.Q0, c.dt, steps, kck, probe, target, params=c.n, wrt=("n",))
        return time.perf_counter() - t0, loss, gn, ctx.last_checkpoint_interval()

    el, loss, gn, k_used = timed(args.kck)
    out = {"value": c.N * steps / el, "unit": "cell-steps/s", "ms_per_gradient": el * 1e3,
           "config": {"workload": "channel_1M_tri_probe_loss_gradient", "cells": c.N, "steps": steps,
                      "checkpoint_every": k_used,
                      "checkpoint_choice": "library (kck = 0)" if args.kck <= 0 else "fixed",
                      "trajectory_GB": (steps + 1) * 3 * c.N * 8 / 1e9 if k_used >= steps else None,
                      "probes": int(probe.size), "wrt": "per-cell Manning n", "dt": c.dt},
           "roofline": gradient_roofline(c.mesh, c.N, steps, k_used, el),
           "loss": loss, "grad_norm": float(np.linalg.norm(gn))}
    el2, loss2, gn2, _ = timed(args.grad_kck)
    out["checkpointed"] = {"value": c.N * steps / el2, "unit": "cell-steps/s", "ms_per_gradient": el2 * 1e3,
                           "checkpoint_every": args.grad_kck,
                           "roofline": gradient_roofline(c.mesh, c.N, steps, args.grad_kck, el2),
                           "bitwise_equal": bool(loss2 == loss and np.array_equal(gn2, gn))}
    return out


def stored_gradient(ctx, c, steps, probe, target):
    """The same gradient with one segment covering the run (checkpoint
    interval = steps): the forward sweep keeps every state, nothing is
    recomputed (bitwise the checkpointed gradient; (steps + 1) 3N doubles of
    device memory)."""
    ctx.grad_probes(c.Q0, c.dt, steps, steps, probe, target, params=c.n, wrt=("n",))  # warm-up (allocates)
    t0 = time.perf_counter()
    ctx.grad_probes(c.Q0, c.dt, steps, steps, probe, target, params=c.n, wrt=("n",))
    el = time.perf_counter() - t0
    return {"value": c.N * steps / el, "unit": "cell-steps/s", "ms_per_gradient": el * 1e3,
            "checkpoint_every": steps, "trajectory_GB": (steps + 1) * 3 * c.N * 8 / 1e9,
            "roofline": gradient_roofline(c.mesh, c.N, steps, steps, el)}


def measure_wet_dry(args, capi, tile):
    """The forward workload with wet/dry fronts: the same channel generator
    and size with bumps up to 4 m (islands above the 3 m surface; seed 1) at
    CFL 0.3, the step kernel's launch shape of the main leg.  The timed region
    and the per-dispatch kernel time as for the headline, and the dry
    fraction (h at h_small) before and after."""
    from hgswe import cases

    c = cases.channel(args.cells, seed=1, nsteps=args.steps, bump_amp=4.0, cfl=0.3)
    ctx = capi.Context(c)
    ctx.set_tile_config(*tile)
    ctx.upload(c.Q0)
    ctx.advance(c.dt, max(args.warmup, 16))
    ctx.prepare(c.dt)
    t0 = time.perf_counter()
    ms_region = ctx.advance_region(c.dt, args.steps)
    el = time.perf_counter() - t0
    ms_dispatch = ctx.advance_timed(c.dt, args.steps)
    q = ctx.download()

    def dry(Q):
        h = Q[: c.N] + c.hstill
        return float(np.mean(h <= c.h_small))

    bpc = bytes_per_cell_step(c.mesh)
    return {"value": c.N * args.steps / el, "unit": "cell-steps/s", "ms_per_step": el * 1e3 / args.steps,
            "region_ms_per_step": ms_region, "kernel_ms": ms_dispatch,
            "frac": bpc * c.N / (ms_region * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "frac_kernel": bpc * c.N / (ms_dispatch * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "config": {"workload": "channel_1M_tri_islands_forward_euler", "cells": c.N, "bump_amplitude_m": 4.0,
                       "cfl": 0.3, "dt": c.dt, "cells_above_surface": float(np.mean(c.zb > 3.0)),
                       "dry_fraction_start": dry(c.Q0), "dry_fraction_end": dry(q),
                       "steps_before_end": args.warmup + 2 * args.steps}}


def batch_instances(n_inst, seed=7, lo=5_000, hi=500_000):
    """64 independent channel instances with N_i log-uniform in [5k, 500k]
    triangles, each with its own bed (seeded bumps), zone n and inlet
    discharge Q ~ U[50, 1000] m^3/s."""
    from hgswe import cases

    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), n_inst)).astype(int)
    Qs = rng.uniform(50.0, 1000.0, n_inst)
    return [cases.channel(int(n), seed=1000 + r, Q_in=float(q), nsteps=100) for r, (n, q) in enumerate(zip(sizes, Qs))]


def measure_batch(args, capi, single_value):
    from hgswe import cases

    t0 = time.perf_counter()
    inst = batch_instances(args.batch)
    B = cases.concat(inst)
    print(f"# batch: {len(inst)} instances, N={B.N}, inlets={B.mesh.n_inlet}, dt={B.dt:.6g} "
          f"(built in {time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    ctx = capi.Context(B)
    ctx.upload(B.Q0)
    T, x, _ = ctx.autotune(B.dt, 100)
    ctx.advance(B.dt, args.warmup)
    t0 = time.perf_counter()
    ctx.advance(B.dt, args.batch_steps)
    el = time.perf_counter() - t0
    out = {"value": B.N * args.batch_steps / el, "unit": "cell-steps/s", "ms_per_step": el * 1e3 / args.batch_steps,
           "single_instance_value": single_value,
           "config": {"workload": f"batch{len(inst)}_channels_5k_500k_forward_euler", "instances": len(inst),
                      "cells": B.N, "steps": args.batch_steps, "dt": B.dt, "tile_cells": T, "xcd_chunk": x,
                      "working_set_MB": round(bytes_per_cell_step(B.mesh) * B.N / 1e6, 1)}}
    # gradient of the summed per-instance probe loss w.r.t. n, zb and every inlet Q
    rng = np.random.default_rng(11)
    off = B.extra["cell_off"]
    probe = np.concatenate([off[r] + rng.choice(c.N, size=200, replace=False) for r, c in enumerate(inst)])
    probe = probe.astype(np.int32)
    QN = ctx.step(B.Q0, B.dt, args.batch_grad_steps)
    target = QN[probe] + B.hstill[probe] + B.zb[probe] + rng.normal(0.0, 0.02, probe.size)
    ctx.grad_probes(B.Q0, B.dt, args.batch_grad_steps, args.kck, probe, target)  # warm-up (allocates)
    t0 = time.perf_counter()
    ctx.grad_probes(B.Q0, B.dt, args.batch_grad_steps, args.kck, probe, target)
    el = time.perf_counter() - t0
    out["gradient"] = {"value": B.N * args.batch_grad_steps / el, "unit": "cell-steps/s",
                       "ms_per_gradient": el * 1e3, "steps": args.batch_grad_steps,
                       "checkpoint_every": ctx.last_checkpoint_interval(), "probes": int(probe.size),
                       "wrt": "n, zb, inlet Q"}
    return out


def measure_ude(args, capi):
    """UDE closure workload: ~50k-triangle channel, n = 0.01 + softplus(MLP(h,
    |u|, zb)) in f32 on the matrix cores every step; 32 perturbed weight sets
    advanced together as a batch, forward and in one reverse sweep that gives
    every set's weight gradient; gradient of the probe loss w.r.t. the 353
    weights of one set."""
    from hgswe import cases, ude

    c = cases.channel(args.ude_cells, seed=4, nsteps=args.ude_steps)
    W = ude.xavier_weights(0)
    ctx = capi.Context(c)
    ctx.set_ude(W, precision=32)
    ctx.upload(c.Q0)
    ctx.advance(c.dt, args.warmup)
    t0 = time.perf_counter()
    ctx.advance(c.dt, args.ude_steps)
    el = time.perf_counter() - t0
    out = {"value": c.N * args.ude_steps / el, "unit": "cell-steps/s", "ms_per_step": el * 1e3 / args.ude_steps,
           "config": {"workload": "channel_50k_tri_ude_closure_f32", "cells": c.N, "steps": args.ude_steps,
                      "dt": c.dt, "network": "3-16-16-1 tanh, 353 weights, f32 matrix cores"}}
    # 32 weight perturbations batched, one weight set per copy of the mesh:
    # as a batch of 32 copies (the mesh stored 32 times) and as 32 fields of
    # one mesh (stored once); both give the same bits
    nset = args.ude_sets
    B = cases.concat([c] * nset)
    Ws = ude.perturbed_sets(W, nset, scale=0.05)
    set_of_cell = np.repeat(np.arange(nset), c.N).astype(np.int32)
    rng = np.random.default_rng(6)
    bprobe = np.concatenate([r * c.N + rng.choice(c.N, size=200, replace=False) for r in range(nset)])
    bprobe = bprobe.astype(np.int32)
    noise = rng.normal(0.0, 0.02, bprobe.size)
    for key, share in (("batched_sets", False), ("batched_sets_shared_mesh", True)):
        bctx = capi.Context(B, share_mesh=share)
        bctx.set_ude(Ws, set_of_cell=set_of_cell, precision=32)
        bctx.upload(B.Q0)
        bctx.advance(B.dt, max(args.warmup, 100))
        t0 = time.perf_counter()
        bctx.advance(B.dt, args.ude_steps)
        el = time.perf_counter() - t0
        out[key] = {"value": B.N * args.ude_steps / el, "unit": "cell-steps/s", "sets": nset, "cells": B.N,
                    "ms_per_step": el * 1e3 / args.ude_steps,
                    "bytes_per_cell_step": bytes_per_cell_step(c.mesh, nset if share else 1)}
        # one reverse sweep over the batch: the gradient of each set's own
        # probe loss (200 probes per instance) w.r.t. its 353 weights
        QB = bctx.step(B.Q0, B.dt, args.ude_steps)
        btarget = QB[bprobe] + B.hstill[bprobe] + B.zb[bprobe] + noise
        bctx.grad_probes(B.Q0, B.dt, args.ude_steps, args.kck, bprobe, btarget, wrt=())  # warm-up (allocates)
        t0 = time.perf_counter()
        bctx.grad_probes(B.Q0, B.dt, args.ude_steps, args.kck, bprobe, btarget, wrt=())  # the weights only
        gws = bctx.ude_grad()
        el = time.perf_counter() - t0
        k_used = bctx.last_checkpoint_interval()
        del bctx
        out[key]["gradient"] = {
            "value": B.N * args.ude_steps / el, "unit": "cell-steps/s", "ms_per_gradient": el * 1e3,
            "steps": args.ude_steps, "checkpoint_every": k_used, "probes": int(bprobe.size),
            "wrt": f"{nset} x 353 closure weights (each set's own instance loss)",
            "grad_norm_min": float(np.min(np.linalg.norm(gws.reshape(nset, -1), axis=1)))}
    # gradient w.r.t. the weights
    rng = np.random.default_rng(5)
    probe = rng.choice(c.N, size=200, replace=False).astype(np.int32)
    QN = ctx.step(c.Q0, c.dt, args.ude_steps)
    target = QN[probe] + c.hstill[probe] + c.zb[probe] + rng.normal(0.0, 0.02, probe.size)
    ctx.grad_probes(c.Q0, c.dt, args.ude_steps, args.kck, probe, target, wrt=())  # warm-up (allocates)
    t0 = time.perf_counter()
    ctx.grad_probes(c.Q0, c.dt, args.ude_steps, args.kck, probe, target, wrt=())  # the weights only
    gw = ctx.ude_grad()
    el = time.perf_counter() - t0
    out["gradient"] = {"value": c.N * args.ude_steps / el, "unit": "cell-steps/s", "ms_per_gradient": el * 1e3,
                       "steps": args.ude_steps, "checkpoint_every": ctx.last_checkpoint_interval(),
                       "wrt": "353 closure weights",
                       "grad_norm": float(np.linalg.norm(gw))}
    return out


def measure_fields(args, capi):
    """K Manning-n fields (lognormal, mean 0.035, sigma 0.3) on the ~50k-
    triangle channel: the K fields as a batch of K copies of the mesh and as
    K fields of one mesh (hg_fields_create), forward steps and the probe-loss
    gradient of every field w.r.t. its own n."""
    from hgswe import cases

    c = cases.channel(args.ude_cells, seed=4, nsteps=args.ude_steps)
    K = args.fields
    rng = np.random.default_rng(31)
    inst = [c.copy_fields(n=np.exp(rng.normal(np.log(0.035) - 0.045, 0.3, c.N))) for _ in range(K)]
    B = cases.concat(inst)
    probe = np.concatenate([f * c.N + rng.choice(c.N, size=200, replace=False) for f in range(K)]).astype(np.int32)
    noise = rng.normal(0.0, 0.02, probe.size)
    out = {"config": {"workload": f"channel_50k_tri_{K}_manning_n_fields", "fields": K, "cells_per_field": c.N,
                      "steps": args.ude_steps, "dt": B.dt}}
    for key, share in (("batch_of_copies", False), ("shared_mesh", True)):
        k = capi.Context(B, share_mesh=share)
        k.upload(B.Q0)
        k.advance(B.dt, max(args.warmup, 100))
        t0 = time.perf_counter()
        k.advance(B.dt, args.ude_steps)
        el = time.perf_counter() - t0
        ms_k = k.advance_timed(B.dt, 50)
        bpc = bytes_per_cell_step(c.mesh, K if share else 1)
        r = {"value": B.N * args.ude_steps / el, "unit": "cell-steps/s", "ms_per_step": el * 1e3 / args.ude_steps,
             "kernel_ms": ms_k, "bytes_per_cell_step": bpc,
             "achieved_GBps": bpc * B.N / (ms_k * 1e-3) / 1e9}
        QN = k.step(B.Q0, B.dt, args.ude_steps)
        target = QN[probe] + B.hstill[probe] + B.zb[probe] + noise
        k.grad_probes(B.Q0, B.dt, args.ude_steps, args.kck, probe, target, wrt=("n",))  # warm-up (allocates)
        t0 = time.perf_counter()
        k.grad_probes(B.Q0, B.dt, args.ude_steps, args.kck, probe, target, wrt=("n",))
        el = time.perf_counter() - t0
        r["gradient"] = {"value": B.N * args.ude_steps / el, "unit": "cell-steps/s", "ms_per_gradient": el * 1e3,
                         "checkpoint_every": k.last_checkpoint_interval(), "wrt": "each field's own per-cell n"}
        out[key] = r
        del k
    return out


def measure_paths(args, capi, case, tile):
    """The other device paths on the forward workload's 1M-triangle mesh, each
    timed as the difference of two runs (n and 2n steps) so that the host
    transfers at the call's ends cancel:
      - adaptive Tsit5 (hg_tsit5_solve, abstol 1e-6, reltol 1e-3): 6 RHS
        evaluations, 7 stage combinations and the error norm per accepted step;
        algorithmic bytes per cell and accepted step 6 x 180 (RHS) + 42 x 24
        (the combinations read y and 1..6 stage derivatives and write 3N each,
        the error norm reads y, y_new and 7 derivatives);
      - forward mode (hg_jvp): the Euler step on dual numbers, one tangent;
      - the step with the reference's UDE network (hg_set_lux_ude: ManningN_h,
        hidden layers [3, 3], tanh, the bump example's bounds) re-evaluating n
        per cell before every step;
      - the step with the variable-n power law re-evaluated before every step."""
    from hgswe import cases  # noqa: F401  (the case is the forward leg's)

    c = case
    out = {}
    n = 20

    def diff(run, k):
        run(k)  # warm-up
        t0 = time.perf_counter()
        a = run(k)
        t1 = time.perf_counter()
        b = run(2 * k)
        t2 = time.perf_counter()
        return (t2 - t1) - (t1 - t0), a, b

    ctx = capi.Context(c)
    ctx.set_tile_config(*tile)
    span = 100 * c.dt  # about 15 accepted Tsit5 steps, 30 for the doubled span

    def tsit(k):
        _, st = ctx.tsit5(c.Q0, 0.0, span * k / n, 1, abstol=1e-6, reltol=1e-3, dt0=c.dt)
        return st
    el, sa, sb = diff(tsit, n)
    acc = sb[0] - sa[0]
    nrhs = sb[2] - sa[2]
    bpc = 6 * bytes_per_cell_step(c.mesh) + 42 * 24
    out["tsit5"] = {"ms_per_accepted_step": el * 1e3 / acc, "accepted_steps": [sa[0], sb[0]],
                    "rejected_steps": [sa[1], sb[1]], "cell_rhs_per_s": c.N * nrhs / el,
                    "bytes_per_cell_step": bpc, "achieved_GBps": bpc * c.N * acc / el / 1e9,
                    "frac": bpc * c.N * acc / el / 1e9 / HBM_PEAK_GBS,
                    "config": {"span_s": [span, 2 * span], "abstol": 1e-6, "reltol": 1e-3}}
    rng = np.random.default_rng(12)
    dQ = rng.normal(size=3 * c.N) * 1e-3
    el, _, _ = diff(lambda k: ctx.jvp(c.Q0, c.dt, 5 * k, dQ0=dQ), n)
    out["jvp"] = {"value": c.N * 5 * n / el, "unit": "cell-steps/s", "ms_per_step": el * 1e3 / (5 * n)}
    del ctx

    ctx = capi.Context(c)
    ctx.set_tile_config(*tile)
    net = capi.lux_net("ManningN_h", [3, 3], ["tanh", "tanh"], (0.03, 0.06), (0.1, 0.5))
    w = rng.normal(size=capi.lux_n_params(net)) * 0.5
    ctx.set_lux_ude(net, w)
    el, _, _ = diff(lambda k: ctx.step(c.Q0, c.dt, 5 * k), n)
    out["lux_network_step"] = {"value": c.N * 5 * n / el, "unit": "cell-steps/s", "ms_per_step": el * 1e3 / (5 * n),
                               "network": "ManningN_h, hidden [3, 3], tanh, FP64"}
    ctx.set_lux_ude(None)
    ctx.set_n_law("power_law", n_lower=0.02, n_upper=0.06, k=0.3, h_mid=1.0)
    el, _, _ = diff(lambda k: ctx.step(c.Q0, c.dt, 5 * k), n)
    out["n_law_step"] = {"value": c.N * 5 * n / el, "unit": "cell-steps/s", "ms_per_step": el * 1e3 / (5 * n),
                         "law": "power law"}
    del ctx
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=1000,
                    help="untimed steps before the timed region (the device clocks settle over several hundred)")
    ap.add_argument("--cells", type=int, default=1_000_000)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--grad-cells", type=int, default=100_000)
    ap.add_argument("--grad-steps", type=int, default=2000)
    ap.add_argument("--grad-kck", type=int, default=50,
                    help="checkpoint interval of the 100k-cell gradient (its stated 50) and of the 1M-cell one beside it")
    ap.add_argument("--kck", type=int, default=0,
                    help="checkpoint interval of the 1M-cell, batch, closure and fields gradients (0: the library's choice)")
    ap.add_argument("--grad-1m-steps", type=int, default=1000,
                    help="steps of the gradient on the forward workload's 1M-triangle mesh")
    ap.add_argument("--no-grad", action="store_true")
    ap.add_argument("--no-wet-dry", action="store_true", help="skip the forward run on the bed with islands")
    ap.add_argument("--batch", type=int, default=64, help="instances of the batched workload (0: skip it)")
    ap.add_argument("--batch-steps", type=int, default=400)
    ap.add_argument("--batch-grad-steps", type=int, default=200)
    ap.add_argument("--ude-cells", type=int, default=50_000)
    ap.add_argument("--ude-steps", type=int, default=1000)
    ap.add_argument("--ude-sets", type=int, default=32)
    ap.add_argument("--no-ude", action="store_true")
    ap.add_argument("--fields", type=int, default=32, help="Manning-n fields of the shared-mesh leg (0: skip it)")
    ap.add_argument("--tile-config", default="", help="T,G: fixed launch shape (tile cells, XCD chunk); no autotune")
    ap.add_argument("--no-paths", action="store_true", help="skip the Tsit5 / forward-mode / network / n-law leg")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shape (256-cell tiles, identity tile mapping)")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="after the timed region, run one device-to-device copy of this many MiB "
                         "(calibrates rocprof FETCH_SIZE / WRITE_SIZE against a kernel of known traffic)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); a run over several GPUs "
              "(domain decomposition with halo exchange) is not provided", file=sys.stderr)
        sys.exit(2)

    from hgswe import capi, cases

    t0 = time.perf_counter()
    case = cases.channel(args.cells, seed=0, nsteps=args.steps)
    print(f"# case {case.name}: N={case.N} faces={case.mesh.n_faces} ghosts={case.G} dt={case.dt:.6g} "
          f"(built in {time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    ctx = capi.Context(case)
    print(f"# {capi.info()}", file=sys.stderr, flush=True)
    ctx.upload(case.Q0)
    if args.tile_config:
        T, G = (int(v) for v in args.tile_config.split(","))
        ctx.set_tile_config(T, G)
    elif not args.no_tune:  # launch shapes are bit-identical; keep the fastest
        T, x, _ = ctx.autotune(case.dt, 200)
        print(f"# autotune: {T}-cell tiles, xcd_chunk={x}", file=sys.stderr, flush=True)
    # the stepping graph for dt and the dispatch events exist before the
    # timed region (the autotuner already built the graph; this covers the
    # fixed-shape paths), so the region holds no host-side graph build
    ctx.prepare(case.dt)
    # the W warm-up steps go through the same entry point as the timed steps,
    # right before them
    if args.warmup > 0:
        ctx.advance_region(case.dt, args.warmup)
    # timed region: K device-resident steps (hipGraph replays of the step
    # kernel alone), synchronized on both sides; HIP events on the context
    # stream bracket the same launches.  ms_region = event time / K is the
    # step kernel's average launch duration over the region, dispatch gaps
    # between the dependent launches included; the roofline uses it.
    t0 = time.perf_counter()
    ms_region = ctx.advance_region(case.dt, args.steps)
    el = time.perf_counter() - t0
    value = case.N * args.steps / el
    # the step kernel alone: the same K steps right after, as single
    # launches with HIP events attached to each dispatch on the context
    # stream (the per-dispatch duration a profiler's kernel trace reports)
    ms_dispatch = ctx.advance_timed(case.dt, args.steps)
    bpc = bytes_per_cell_step(case.mesh)
    achieved = bpc * case.N / (ms_region * 1e-3) / 1e9
    achieved_kernel = bpc * case.N / (ms_dispatch * 1e-3) / 1e9
    out = {
        "metric": "cell-steps/s fwd and fwd+adjoint gradient, 1M-cell tri mesh FP64; % HBM roof",
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "channel_1M_tri_forward_euler", "cells": case.N, "faces": int(case.mesh.n_faces),
                   "ghosts": case.G, "cfl": 0.5, "dt": case.dt, "parallelism": "single-gpu",
                   "tile_cells": ctx.tile_config()[0], "xcd_chunk": ctx.tile_config()[1]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profiles(),
                     "kernel": "k_swe_tile", "bytes_per_cell_step": bpc, "region_ms_per_step": ms_region,
                     "kernel_ms": ms_dispatch, "frac_kernel": achieved_kernel / HBM_PEAK_GBS},
    }
    if args.copy_mb > 0:  # one launch of a copy kernel of known traffic (calibrates rocprof counters)
        nbytes = args.copy_mb * (1 << 20)
        out["copy_probe"] = {"bytes_each_way": nbytes, "GBps": capi.probe_copy(ctx, nbytes, 1)}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(case, args.cpu_seconds)
    tile = ctx.tile_config()
    del ctx
    if not args.no_paths:
        try:  # an auxiliary leg: a failure is reported in the line, not fatal to it
            out["paths"] = measure_paths(args, capi, case, tile)
        except Exception as e:
            out["paths"] = {"error": repr(e)}
    if not args.no_wet_dry:
        out["wet_dry"] = measure_wet_dry(args, capi, tile)
    if not args.no_grad:
        out["gradient_1M"] = measure_gradient_full(args, capi, case, tile)
        out["gradient"] = measure_gradient(args, capi)
    if args.batch > 0:
        out["batch"] = measure_batch(args, capi, value)
    if not args.no_ude:
        out["ude"] = measure_ude(args, capi)
    if args.fields > 0:
        out["fields"] = measure_fields(args, capi)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
