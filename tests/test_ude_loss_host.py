"""The UDE / inversion loss terms (hgswe.inversion.final_state_loss) against
the reference's own UDE training run (swe_2D_UDE.jl:522-599 as run on the
n(h) bump truth): every one of its 100 loss values from the predicted states
it saved (iterations 0, 1, 25, 50, 99 here), WSE and uv parts and total."""
import os

import numpy as np

from conftest import GOLDEN, load_golden_case
from hgswe import inversion as inv


def test_ude_loss_matches_reference_history():
    c = load_golden_case("bump", ic=(0.33, 0.2, 0.0, 0.0))
    t = np.load(os.path.join(GOLDEN, "bump_nh", "truth.npz"))
    r = np.load(os.path.join(GOLDEN, "ude_bump", "history.npz"))
    truth = {"WSE_truth": t["wse_truth"], "u_truth": t["u_truth"], "v_truth": t["v_truth"]}
    for j, k in enumerate(r["iterations"]):
        h, u, v = r["inverted_hs"][j], r["inverted_us"][j], r["inverted_vs"][j]
        QN = np.concatenate([h - c.hstill, u * h, v * h])  # the saved velocities are q / h (swe_2D_UDE.jl:408-409)
        total, parts, *_ = inv.final_state_loss(c, QN, truth, inv.ude_settings(), zb=t["zb_cell_truth"])
        for got, key in ((parts["wse"], "loss_pred_WSEs"), (parts["uv"], "loss_pred_uvs"), (total, "loss_totals")):
            assert abs(got / r[key][k] - 1.0) < 1e-13, (k, key)


def test_ude_loss_gradient_matches_differences():
    c = load_golden_case("bump", ic=(0.33, 0.2, 0.0, 0.0))
    t = np.load(os.path.join(GOLDEN, "bump_nh", "truth.npz"))
    r = np.load(os.path.join(GOLDEN, "ude_bump", "history.npz"))
    truth = {"WSE_truth": t["wse_truth"], "u_truth": t["u_truth"], "v_truth": t["v_truth"]}
    h, u, v = r["inverted_hs"][1], r["inverted_us"][1], r["inverted_vs"][1]
    QN = np.concatenate([h - c.hstill, u * h, v * h])
    _, _, g, _, _ = inv.final_state_loss(c, QN, truth, inv.ude_settings(), zb=t["zb_cell_truth"])
    rng = np.random.default_rng(0)
    d = rng.normal(size=QN.size) * 1e-3
    f = lambda q: inv.final_state_loss(c, q, truth, inv.ude_settings(), zb=t["zb_cell_truth"])[0]
    e = 1e-4
    fd = (f(QN + e * d) - f(QN - e * d)) / (2 * e)
    assert abs(fd - g @ d) < 1e-7 * abs(fd)
