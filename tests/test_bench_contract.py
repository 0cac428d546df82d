"""The benchmark's contract, checked on the CPU: the single-GPU guard, the
algorithmic bytes per cell-step (96 + 28 x degree: 180 B for triangles, 208 B
for quads), the HBM traffic figure read from the committed rocprof summary,
and the keys of the CPU-baseline objects."""
import importlib.util
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def bench_module():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_algorithmic_bytes_per_cell_step():
    from hgswe import cases

    b = bench_module()
    tri = cases.channel(2_000, seed=1, Lx=400.0, Ly=40.0, nsteps=1)
    quad = cases.dam_break(nx=20, ny=4)
    assert b.bytes_per_cell_step(quad.mesh) == 24 + 24 + 48 + 28 * 4 == 208
    deg = np.diff(tri.mesh.face_off)
    assert np.all(deg == 3)
    assert b.bytes_per_cell_step(tri.mesh) == 180


def test_traffic_figure_comes_from_the_committed_profile():
    b = bench_module()
    t = b.traffic_from_profiles()
    line = [s for s in open(os.path.join(ROOT, "profiles", "k_swe_tile_channel1M_traffic.txt"))
            if s.startswith("traffic_bytes_per_launch:")]
    assert len(line) == 1 and t == float(line[0].split(":")[1])
    assert 1.0e8 < t < 2.5e8  # HBM bytes per launch of the 1M-cell step kernel


def test_cpu_baseline_objects():
    from hgswe import cases

    b = bench_module()
    c = cases.dam_break(nx=20, ny=4)
    out = b.cpu_baseline(c, 0.05)
    assert set(out) >= {"value", "unit", "cores", "kind", "sample", "all_cores"}
    assert out["cores"] == 1 and out["kind"] == "port" and out["value"] > 0
    assert out["all_cores"]["cores"] == b.host_threads() and out["all_cores"]["value"] > 0


def test_gradient_bytes_per_horizon_cell_step():
    """Forward sweep + recompute (180 B each on triangles), the reverse step
    (24 + 24 + 24 + 48 + 28 x 3 + 16 = 220 B) and the checkpoint write 24/k."""
    from hgswe import cases

    b = bench_module()
    tri = cases.channel(2_000, seed=1, Lx=400.0, Ly=40.0, nsteps=1)
    assert b.bytes_per_gradient_cell_step(tri.mesh, 50) == 180 + 180 + 220 + 24 / 50
    r = b.gradient_roofline(tri.mesh, 1_000_000, 1000, 50, 0.1)
    assert abs(r["achieved"] - 580.48 * 1e9 / 0.1 / 1e9) < 1e-6 and r["frac"] == r["achieved"] / 8000.0
    # one segment covering the run: no recompute, every state stored (24 B)
    assert b.bytes_per_gradient_cell_step(tri.mesh, 1000, 1000) == 180 + 24 + 220
    assert b.gradient_roofline(tri.mesh, 1_000_000, 1000, 1000, 0.1)["bytes_per_horizon_cell_step"] == 424


def test_shared_mesh_bytes_per_cell_step():
    """K fields of one mesh: each field's state (48 B) and n (8 B), the mesh
    data (h_still, zb, S0x, S0y, area: 40 B; faces 28 B each) once for all."""
    from hgswe import cases

    b = bench_module()
    tri = cases.channel(2_000, seed=1, Lx=400.0, Ly=40.0, nsteps=1)
    assert b.bytes_per_cell_step(tri.mesh, 1) == 180
    assert b.bytes_per_cell_step(tri.mesh, 32) == 56 + (40 + 84) / 32
