"""SRH-2D case files (.srhgeom / .srhhydro / .srhmat): reader and writer.

Reader semantics follow Hydrograd.jl's parsers:
  * .srhgeom  -- src/utilities/SRH_2D/SRH_2D_SRHGeom.jl:150-223 (Elem, Node,
    NodeString with continuation lines, bed elevation = node mean)
  * .srhhydro -- src/utilities/SRH_2D/SRH_2D_SRHHydro.jl:25-93 (BC, IQParams,
    EWSParamsC, ManningsN keyed by integer id)
  * .srhmat   -- src/utilities/SRH_2D/SRH_2D_SRHMat.jl:30-108 (Material lists;
    cells in no material get zone 0, process_SRH_2D_input.jl:136-153)
The writer emits the keys process_SRH_2D_input.jl:55-69 requires plus at
least one BC line, so a generated case also loads in the reference.
"""
from __future__ import annotations

import os
import shlex

import numpy as np


def read_srhgeom(path):
    elems = {}
    nodes = {}
    nodestrings = {}
    cur = None
    with open(path) as fh:
        for line in fh:
            parts = line.split()
            if not parts:
                continue
            tag = parts[0]
            if tag == "Elem":
                elems[int(parts[1])] = [int(v) for v in parts[2:]]
                cur = None
            elif tag == "Node":
                nodes[int(parts[1])] = [float(v) for v in parts[2:5]]
                cur = None
            elif tag == "NodeString":
                cur = int(parts[1])
                nodestrings[cur] = [int(v) for v in parts[2:]]
            elif tag.lower() in ("name", "gridunit") or "srhgeom" in tag.lower():
                cur = None
            elif cur is not None:
                nodestrings[cur].extend(int(v) for v in parts)
    n_nodes = max(nodes)
    xyz = np.zeros((n_nodes, 3))
    for k, v in nodes.items():
        xyz[k - 1, : len(v)] = v
    n_cells = max(elems)
    maxdeg = max(len(v) for v in elems.values())
    cells = np.full((n_cells, maxdeg), -1, dtype=np.int64)
    for k, v in elems.items():
        cells[k - 1, : len(v)] = np.asarray(v) - 1
    ns = {k: [x - 1 for x in v] for k, v in nodestrings.items()}
    return xyz, cells, ns


def read_srhhydro(path):
    out = {"BC": {}, "IQParams": {}, "EWSParamsC": {}, "ManningsN": {}}
    with open(path) as fh:
        for line in fh:
            line = line.strip()
            if not line or line.startswith("#"):
                continue
            parts = shlex.split(line)
            key = parts[0]
            if key in ("BC", "IQParams", "EWSParamsC", "ManningsN"):
                out[key][int(parts[1])] = parts[2:]
            else:
                out[key] = parts[1:]
    return out


def read_srhmat(path, n_cells):
    zone = np.zeros(n_cells, dtype=np.int32)
    cur = 0
    names = {}
    with open(path) as fh:
        for line in fh:
            parts = line.split()
            if not parts:
                continue
            if parts[0] == "NMaterials":
                continue
            if parts[0] == "MatName":
                names[int(parts[1])] = " ".join(parts[2:]).strip('"')
                continue
            if parts[0] == "Material":
                cur = int(parts[1])
                ids = parts[2:]
            elif parts[0].lstrip("-").isdigit() and cur > 0:
                ids = parts
            else:
                continue
            for v in ids:
                zone[int(v) - 1] = cur
    return zone, names


def write_case(dirpath, name, node_xyz, cells, nodestrings, bcs, iq=None, ews=None, manning=None, zone=None):
    """Write <name>.srhgeom/.srhhydro/.srhmat.  ``bcs`` maps boundary id ->
    type string, ``iq``/``ews`` map id -> value, ``manning`` maps zone -> n."""
    os.makedirs(dirpath, exist_ok=True)
    with open(os.path.join(dirpath, name + ".srhgeom"), "w") as fh:
        fh.write("SRHGEOM 30\nName \"generated\"\n\nGridUnit \"Meters\"\n")
        for i, row in enumerate(cells):
            ids = " ".join(str(int(v) + 1) for v in row if v >= 0)
            fh.write(f"Elem {i + 1} {ids}\n")
        for i, (x, y, z) in enumerate(node_xyz):
            fh.write(f"Node {i + 1} {float(x)!r} {float(y)!r} {float(z)!r}\n")
        for k, ns in sorted(nodestrings.items()):
            fh.write(f"NodeString {k} " + " ".join(str(int(v) + 1) for v in ns) + "\n")
    with open(os.path.join(dirpath, name + ".srhhydro"), "w") as fh:
        fh.write(f"SRHHYDRO 30\nCase \"{name}\"\nDescription \"generated\"\nRunType FLOW\nModelTemp OFF\n"
                 "UnsteadyOutput UNSTEADY\nSimTime 0.0 0.1 1\nTurbulenceModel PARABOLIC\n"
                 "ParabolicTurbulence 0.7\nInitCondOption DRY\n"
                 f"Grid \"{name}.srhgeom\"\nHydroMat \"{name}.srhmat\"\nOutputFormat XMDFC SI\n"
                 "OutputOption 1\nOutputInterval 0.1\n")
        for z, n in sorted((manning or {0: 0.03}).items()):
            fh.write(f"ManningsN {int(z)} {float(n)!r}\n")
        for k, t in sorted(bcs.items()):
            fh.write(f"BC {k} {t}\n")
        for k, q in sorted((iq or {}).items()):
            fh.write(f"IQParams {k} {float(q)!r} SI CONVEYANCE\n")
        for k, w in sorted((ews or {}).items()):
            fh.write(f"EWSParamsC {k} {float(w)!r} SI C\n")
    zone = np.zeros(len(cells), dtype=np.int32) if zone is None else zone
    with open(os.path.join(dirpath, name + ".srhmat"), "w") as fh:
        mats = sorted(set(int(z) for z in zone if z > 0))
        fh.write(f"SRHMAT 30\nNMaterials {max(mats) if mats else 1}\n")
        for m in mats:
            fh.write(f"MatName {m} \"zone{m}\"\n")
            ids = np.nonzero(zone == m)[0] + 1
            fh.write(f"Material {m} " + " ".join(str(int(v)) for v in ids) + "\n")
