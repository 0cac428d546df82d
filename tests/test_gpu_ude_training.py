"""UDE training on the device: the reference's ManningN_h example
(UDE/ManningN/oneD_channel_with_bump/training: network [3, 3] tanh, h bounds
0.1..0.5, n bounds 0.03..0.06, Adam at lr 0.1) against the n(h) bump truth,
with the reference's UDE loss on the final state after 100 s of Euler steps
and the gradient through the device adjoint into the network weights.

The reference's run (Tsit5 forward, ForwardDiff gradients, Lux's own random
initial weights, which cannot be reproduced without Lux) reached a loss of
8.5e-6 after 40 Adam iterations; this run starts from its own Glorot-uniform
weights, so the bar is the same order of magnitude, a learned n(h) close to
the truth n(h) on the final state, and bitwise repeatable training.
"""
import json
import os
import sys

import numpy as np
import pytest

import lux_ude as lx
from conftest import GOLDEN, load_golden_case
from hgswe import capi, cases
from hgswe import inversion as inv

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def setup():
    rc = json.load(open(os.path.join(GOLDEN, "ude_bump", "run_control.json")))["UDE_options"]
    cfg = rc["UDE_NN_config"]
    c = load_golden_case("bump", ic=tuple(rc["UDE_forward_simulation_initial_condition_constant_values"]))
    t = np.load(os.path.join(GOLDEN, "bump_nh", "truth.npz"))
    truth = {"WSE_truth": t["wse_truth"], "u_truth": t["u_truth"], "v_truth": t["v_truth"],
             "zb_cell_truth": t["zb_cell_truth"]}
    acts = [capi.ACTIVATIONS[a] for a in cfg["activations"]]
    net = lx.Net(lx.MANNING_H, cfg["hidden_layers"], acts, tuple(cfg["output_bounds"]), tuple(cfg["h_bounds"]))
    # Glorot-uniform Dense weights, zero biases, LayerNorm scale 1 / bias 0 (Lux's initialisers)
    rng = np.random.default_rng(1111)
    offs, out, P = net.layout()
    P0 = np.zeros(P)
    for o in offs:
        w, win = o["w"], o["in"]
        a = np.sqrt(6.0 / (w + win))
        P0[o["W"]:o["W"] + w * win] = rng.uniform(-a, a, w * win)
        P0[o["g"]:o["g"] + w] = 1.0
    a = np.sqrt(6.0 / (out["in"] + 1))
    P0[out["W"]:out["W"] + out["in"]] = rng.uniform(-a, a, out["in"])
    dnet = capi.lux_net("ManningN_h", cfg["hidden_layers"], cfg["activations"], cfg["output_bounds"], cfg["h_bounds"])
    lr = float(rc["UDE_learning_rates"][0])
    return c, t, truth, dnet, P0, lr


def test_ude_training_learns_n_of_h():
    c, t, truth, dnet, P0, lr = setup()
    dt = 0.02  # CFL 0.65 on this mesh
    nsteps = int(round(100.0 / dt))
    ctx = capi.Context(c)
    ctx.set_lux_ude(dnet, P0)
    res = inv.train_ude(ctx, c, truth, P0, dt, nsteps, kck=100, lr=lr, iters=40,
                        callback=lambda it, L, p: say(f"  it {it}: loss {L:.4e}") if it % 5 == 0 else None)
    QN = ctx.step(c.Q0, dt, nsteps, params=res.params)
    n = ctx.lux_eval(QN)
    err = np.abs(n - t["ManningN_cells_truth"])
    ref = np.load(os.path.join(GOLDEN, "ude_bump", "history.npz"))["loss_totals"]
    say(f"  loss {res.loss[0]:.3e} -> {res.loss[-1]:.3e} (reference run at iteration 39: {ref[39]:.3e}); "
        f"|n - n_truth| mean {err.mean():.2e} max {err.max():.2e}")
    assert res.loss[-1] < res.loss[0] / 300
    assert res.loss[-1] < 5 * ref[39]
    assert err.mean() < 2e-3
    # bitwise repeatable: the same first iterations from a fresh context
    again = capi.Context(c)
    again.set_lux_ude(dnet, P0)
    r2 = inv.train_ude(again, c, truth, P0, dt, nsteps, kck=100, lr=lr, iters=3)
    assert r2.loss == res.loss[:3]
