"""SRH-2D writer: generated cases written to .srhgeom / .srhhydro / .srhmat,
read back through the same readers the reference's example cases go through,
and rebuilt bitwise (CPU).

The written .srhhydro must also satisfy the reference's own reader:
parse_srhhydro_file (src/utilities/SRH_2D/SRH_2D_SRHHydro.jl:25-84) splits
each line on whitespace and converts the values of SimTime (3 floats),
ParabolicTurbulence, OutputOption (Int), OutputInterval, ManningsN (Int id,
Float64 n), IQParams / EWSParamsC (3 fields after the id); BC lines are kept
only if at least one exists (:83, a case without one raises KeyError on
srhhydro_content["BC"]); process_SRH_2D_input (src/utilities/
process_SRH_2D_input.jl:40-69) then reads the keys listed in REQUIRED.
"""
import json
import os

import numpy as np
import pytest

from hgswe import cases, srh2d

REQUIRED = ["ManningsN", "BC", "Grid", "RunType", "TurbulenceModel", "OutputFormat", "OutputInterval", "Case",
            "ParabolicTurbulence", "Description", "OutputOption", "HydroMat", "InitCondOption", "SimTime",
            "SRHHYDRO", "ModelTemp", "UnsteadyOutput"]


def _reference_style_parse(path):
    """The line rules of parse_srhhydro_file, restated: whitespace split,
    lines of one token skipped, typed values where the reference parses
    them (a value that does not parse raises, as Julia's parse would)."""
    content, mann, bc, iq, ews = {}, {}, {}, {}, {}
    for line in open(path):
        parts = line.strip().split()
        if len(parts) <= 1:
            continue
        k = parts[0]
        if k == "ManningsN":
            mann[int(parts[1])] = float(parts[2])
        elif k == "BC":
            bc[int(parts[1])] = parts[2]
        elif k == "IQParams":
            assert len(parts) >= 5
            iq[int(parts[1])] = parts[2:5]
        elif k == "EWSParamsC":
            assert len(parts) >= 5
            ews[int(parts[1])] = parts[2:5]
        elif k == "OutputFormat":
            content[k] = parts[1:3]
        elif k == "SimTime":
            content[k] = [float(x) for x in parts[1:4]]
            assert len(content[k]) == 3
        elif k in ("ParabolicTurbulence", "OutputInterval"):
            content[k] = float(parts[1])
        elif k == "OutputOption":
            content[k] = int(parts[1])
        else:
            content[k] = parts[1]
    for key, d in (("ManningsN", mann), ("BC", bc), ("IQParams", iq), ("EWSParamsC", ews)):
        if d:
            content[key] = d
    return content


def _write(case, dirpath, name, manning):
    xyz = np.column_stack([case.mesh.node_xy, case.extra["z_nodes"]])
    bcs = case.extra["bc_types"]
    inlet = sorted(k for k, t in bcs.items() if t.upper() == "INLET-Q")
    exits = sorted(k for k, t in bcs.items() if t.upper() == "EXIT-H")
    srh2d.write_case(dirpath, name, xyz, case.mesh.cell_nodes, case.extra["nodestrings"], bcs,
                     iq={k: float(q) for k, q in zip(inlet, case.inletQ)},
                     ews={k: float(w) for k, w in zip(exits, case.exitWSE)}, manning=manning, zone=case.zone)
    N = case.N
    ic = os.path.join(dirpath, "ic.json")
    wstill = case.zb + case.hstill
    with open(ic, "w") as fh:
        json.dump({"wse": (case.Q0[:N] + wstill).tolist(), "wstill": wstill.tolist(),
                   "q_x": case.Q0[N:2 * N].tolist(), "q_y": case.Q0[2 * N:].tolist()}, fh)
    return ic


def _same_mesh(a, b):
    for f in ("node_xy", "cell_nodes", "cell_deg", "edge_nodes", "face_off", "face_nbr", "face_edge", "face_nx",
              "face_ny", "face_len", "area", "centroid", "ghost_edge", "ghost_cell", "ghost_slot", "ghost_kind",
              "ghost_bc", "ghost_nx", "ghost_ny", "ghost_len", "inlet_off", "inlet_ghost"):
        x, y = getattr(a, f), getattr(b, f)
        assert x.shape == y.shape and np.array_equal(x, y), f
    assert a.n_inlet == b.n_inlet and a.n_exit == b.n_exit


def _same_case(a, b):
    _same_mesh(a.mesh, b.mesh)
    for f in ("hstill", "hstill_ghost", "zb", "zb_ghost", "S0x", "S0y", "n", "inletQ", "exitWSE", "Q0"):
        x, y = getattr(a, f), getattr(b, f)
        assert x.shape == y.shape and np.array_equal(x, y), f


@pytest.mark.parametrize("which", ["dam_break", "channel5k"])
def test_written_case_reads_back_bitwise(tmp_path, which):
    if which == "dam_break":
        c = cases.dam_break()
        manning = {0: 0.03}
    else:
        c = cases.channel(5_000, seed=9, Lx=2_000.0, Ly=200.0, Q_in=40.0, nsteps=1)
        manning = {int(z): float(v) for z, v in enumerate(c.zone_n)}
    ic = _write(c, str(tmp_path), which, manning)
    r = cases.load_srh_case(str(tmp_path), which + ".srhhydro", ic=ic, dt=c.dt, nsteps=c.nsteps)
    _same_case(c, r)
    assert np.array_equal(r.zone, c.zone)


@pytest.mark.parametrize("which", ["dam_break", "channel5k"])
def test_written_srhhydro_satisfies_the_reference_reader(tmp_path, which):
    c = cases.dam_break() if which == "dam_break" else cases.channel(5_000, seed=9, Lx=2_000.0, Ly=200.0,
                                                                      nsteps=1)
    manning = {0: 0.03} if which == "dam_break" else {int(z): float(v) for z, v in enumerate(c.zone_n)}
    _write(c, str(tmp_path), which, manning)
    content = _reference_style_parse(os.path.join(str(tmp_path), which + ".srhhydro"))
    missing = [k for k in REQUIRED if k not in content]
    assert not missing, missing
    assert len(content["BC"]) >= 1
    assert content["Grid"].strip('"') == which + ".srhgeom" and content["HydroMat"].strip('"') == which + ".srhmat"
    bcs = c.extra["bc_types"]
    assert content["BC"] == {k: t for k, t in bcs.items()}
    n_in = sum(1 for t in bcs.values() if t == "INLET-Q")
    n_ex = sum(1 for t in bcs.values() if t == "EXIT-H")
    assert len(content.get("IQParams", {})) == n_in and len(content.get("EWSParamsC", {})) == n_ex
    for v in content.get("IQParams", {}).values():
        float(v[0])
    for v in content.get("EWSParamsC", {}).values():
        float(v[0])
    assert set(content["ManningsN"]) == set(manning)


def test_material_file_round_trip(tmp_path):
    """Zones 1..5 written as Material lists come back per cell; cells in no
    material read as zone 0 (process_SRH_2D_input.jl:136-153)."""
    c = cases.channel(3_000, seed=4, Lx=1_000.0, Ly=100.0, nsteps=1)
    zone = c.zone.copy()
    zone[::7] = 0
    xyz = np.column_stack([c.mesh.node_xy, c.extra["z_nodes"]])
    srh2d.write_case(str(tmp_path), "m", xyz, c.mesh.cell_nodes, c.extra["nodestrings"], c.extra["bc_types"],
                     zone=zone)
    z, names = srh2d.read_srhmat(os.path.join(str(tmp_path), "m.srhmat"), c.N)
    assert np.array_equal(z, zone)
    assert set(names) == set(int(v) for v in np.unique(zone) if v > 0)
