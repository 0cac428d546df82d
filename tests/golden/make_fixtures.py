"""Copies the reference's example data used as test fixtures into
tests/golden/ (run once, in a container where the Hydrograd.jl checkout is
mounted at /root/reference; the committed outputs are data only).

Sources (paths relative to the Hydrograd.jl repository, MIT licence, see
LICENSE_Hydrograd.txt):
  examples/SWE_2D/forward_simulation/oneD_channel_with_bump/       -> bump/
  examples/SWE_2D/forward_simulation/oneD_channel_uniform_flow/     -> uniform/
  examples/SWE_2D/forward_simulation/Savannah_River/                -> savannah/
  examples/SWE_2D/sensitivity_analysis/ManningN/oneD_channel_with_bump/    -> bump_sens/
  examples/SWE_2D/sensitivity_analysis/ManningN/oneD_channel_uniform_flow/ -> uniform_sens/
    (sensitivity_results.json -> sens.npz, forward_simulation_results.json -> traj.npz)
  examples/SWE_2D/forward_simulation/oneD_channel_with_bump_ManningN_h/     -> bump_nh/
  examples/SWE_2D/forward_simulation/Savannah_River_ManningN_ks_h_Umag/     -> savannah_ks/
    (variable Manning n: the sigmoid n(h) law and the h_Umag_ks law; their
    case files and initial condition equal those of bump/ and savannah/, so
    only run_control.json and the truth are copied)
  examples/SWE_2D/sensitivity_analysis/ManningN/Savana_River/               -> savannah_sens/
    (ForwardDiff's dQ(200 s)/dn_zone on the Savannah mesh; case files and
    initial condition equal those of savannah/: run_control.json, sens.npz)
  examples/SWE_2D/UDE/ManningN/oneD_channel_with_bump/training/             -> ude_bump/
    (the reference's UDE training run on the n(h) bump truth: its 100 loss
    values, WSE and uv parts, and the predicted final states h, u = q_x/h,
    v = q_y/h of iterations 0, 1, 25, 50 and 99; run_control.json)

The SRH-2D case files are copied verbatim; the JSON truth / sensitivity
files (written by JSON3 with shortest round-trip Float64 text, so parsing is
exact) are stored as .npz arrays.  Matrices are kept in the reference's
column-major flattening.
"""
import json
import os
import shutil
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))
EX = os.path.join(REF, "examples", "SWE_2D")


def copy_case(src_dir, dst, names):
    os.makedirs(dst, exist_ok=True)
    for n in names:
        shutil.copyfile(os.path.join(src_dir, n), os.path.join(dst, n))


def json_to_npz(src, dst, keys=None):
    with open(src) as fh:
        d = json.load(fh)
    arrs = {}
    for k, v in d.items():
        if keys is not None and k not in keys:
            continue
        if isinstance(v, str):
            arrs[k] = np.array(v)
            continue
        try:
            arrs[k] = np.asarray(v, dtype=np.float64)
        except (TypeError, ValueError):
            continue
    np.savez_compressed(dst, **arrs)


def main():
    fw = os.path.join(EX, "forward_simulation")
    # bump
    d = os.path.join(fw, "oneD_channel_with_bump")
    base = "oneD_channel_with_bump_refined"
    copy_case(d, os.path.join(OUT, "bump"), [base + ".srhgeom", base + ".srhhydro", base + ".srhmat",
                                             "analytical_solution_from_FullSWOF.dat"])
    json_to_npz(os.path.join(d, "forward_simulation_solution_truth.json"), os.path.join(OUT, "bump", "truth.npz"))
    # uniform flow
    d = os.path.join(fw, "oneD_channel_uniform_flow")
    base = "oneD_channel_uniform_flow_refined"
    copy_case(d, os.path.join(OUT, "uniform"), [base + ".srhgeom", base + ".srhhydro", base + ".srhmat"])
    json_to_npz(os.path.join(d, "forward_simulation_solution_truth.json"), os.path.join(OUT, "uniform", "truth.npz"))
    # Savannah River (mixed triangles / quads, 6 zones, inlet + exit + walls)
    d = os.path.join(fw, "Savannah_River")
    copy_case(d, os.path.join(OUT, "savannah"), ["savana_SI.srhgeom", "savana_SI.srhhydro", "savana_SI.srhmat"])
    json_to_npz(os.path.join(d, "forward_simulation_solution_truth.json"),
                os.path.join(OUT, "savannah", "truth.npz"))
    json_to_npz(os.path.join(d, "forward_simulation_initial_condition.json"),
                os.path.join(OUT, "savannah", "ic.npz"))
    # ForwardDiff sensitivities dQ(t_end)/dn_zone
    sa = os.path.join(EX, "sensitivity_analysis", "ManningN")
    for src, dst, base in (("oneD_channel_with_bump", "bump_sens", "oneD_channel_with_bump_refined"),
                           ("oneD_channel_uniform_flow", "uniform_sens", "oneD_channel_uniform_flow_refined")):
        d = os.path.join(sa, src)
        copy_case(d, os.path.join(OUT, dst), [base + ".srhgeom", base + ".srhhydro", base + ".srhmat",
                                              "run_control.json"])
        json_to_npz(os.path.join(d, "sensitivity_results.json"), os.path.join(OUT, dst, "sens.npz"))
        # 101 saved Tsit5 frames of the forward pass (3N per frame, t = 0, 2, ..., T)
        json_to_npz(os.path.join(d, "forward_simulation_results.json"), os.path.join(OUT, dst, "traj.npz"))
    # variable Manning n laws
    for src, dst in (("oneD_channel_with_bump_ManningN_h", "bump_nh"),
                     ("Savannah_River_ManningN_ks_h_Umag", "savannah_ks")):
        d = os.path.join(fw, src)
        copy_case(d, os.path.join(OUT, dst), ["run_control.json"])
        json_to_npz(os.path.join(d, "forward_simulation_solution_truth.json"), os.path.join(OUT, dst, "truth.npz"))
    # Savannah River sensitivities
    d = os.path.join(sa, "Savana_River")
    copy_case(d, os.path.join(OUT, "savannah_sens"), ["run_control.json"])
    json_to_npz(os.path.join(d, "sensitivity_results.json"), os.path.join(OUT, "savannah_sens", "sens.npz"))
    # UDE training history
    d = os.path.join(EX, "UDE", "ManningN", "oneD_channel_with_bump", "training")
    copy_case(d, os.path.join(OUT, "ude_bump"), ["run_control.json"])
    with open(os.path.join(d, "UDE_results_losses_and_ODE_solution.json")) as fh:
        r = json.load(fh)
    its = np.array([0, 1, 25, 50, 99])
    arrs = {k: np.asarray(r[k], dtype=np.float64) for k in ("loss_totals", "loss_pred_WSEs", "loss_pred_uvs",
                                                           "ManningN_cells", "ManningN_cell_truth")}
    for k in ("inverted_hs", "inverted_us", "inverted_vs"):
        arrs[k] = np.asarray(r[k], dtype=np.float64)[its]
    arrs["iterations"] = its
    np.savez_compressed(os.path.join(OUT, "ude_bump", "history.npz"), **arrs)
    shutil.copyfile(os.path.join(REF, "LICENSE"), os.path.join(OUT, "LICENSE_Hydrograd.txt"))


if __name__ == "__main__":
    main()
