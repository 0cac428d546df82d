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
    shutil.copyfile(os.path.join(REF, "LICENSE"), os.path.join(OUT, "LICENSE_Hydrograd.txt"))


if __name__ == "__main__":
    main()
