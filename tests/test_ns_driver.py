"""The fdm_ns_cube driver (fdm_amd/ns_cube_main.cpp): the reference's Config
semantics (INI file + --section:key=value overrides, ny read from the key
"nx") on the CPU with --dry-run, and on the GPU the VTK snapshots of
velocity_plotter::vtk_out's format against the oracle's NSCube."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "fdm_amd", "fdm_ns_cube")


def driver():
    if not os.path.exists(BIN):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "fdm_amd")])
    return BIN


def test_driver_config_and_overrides(tmp_path):
    ini = tmp_path / "ns_cube.ini"
    ini.write_text("[ns]\nnx = 15 ; interior points per axis\nRe=100\n# comment\nny = 7\n[plot]\nvtk 1\n")
    r = subprocess.run([driver(), "--dry-run", f"--config={ini}", "--ns:dt=0.01", "--plot:interval=5", "--ns:nz=31"],
                       capture_output=True, text=True, timeout=60, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    line1, line2 = r.stdout.splitlines()[:2]
    assert "nx 15 ny 15 nz 31" in line1          # ny reads "nx" (src/ns_cube.h:58); [ns] ny is ignored
    assert "Re 100" in line1 and "dt 0.01" in line1
    assert "interval 5" in line2 and "vtk 1" in line2 and "gpu_precision f64" in line2


def read_vtk(path, nx, ny, nz):
    lines = open(path).read().splitlines()
    assert lines[0] == "# vtk DataFile Version 3.0"
    assert lines[3] == "DATASET STRUCTURED_POINTS"
    assert lines[4] == f"DIMENSIONS {nx} {ny} {nz}"
    i = lines.index("VECTORS u double")
    vals = np.array([[float(t) for t in l.split()] for l in lines[i + 1:i + 1 + nx * ny * nz]])
    return int(lines[1].split()[1]), vals.reshape(nz, ny, nx, 3)


@pytest.mark.gpu
def test_driver_vtk_matches_oracle(tmp_path):
    print("[case] fdm_ns_cube driver 15^3, 20 steps, vtk every 10", flush=True)
    sys.stdout.flush()
    n, steps = 15, 20
    r = subprocess.run([driver(), f"--ns:nx={n}", f"--ns:nz={n}", "--ns:Re=250", "--ns:dt=0.01", f"--ns:steps={steps}",
                        "--plot:vtk=1", "--plot:png=0", "--plot:interval=10"],
                       capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    assert "It took me" in r.stdout
    o = O.NsCube(n, n, n, u0=1.0, Re=250.0, dt=0.01)
    for t in (0, 10, 20):
        if t:
            o.step(10)
        step, vel = read_vtk(tmp_path / f"step_{t:07d}.vtk", n, n, n)
        assert step == t
        u, v, w = o.get("u"), o.get("v"), o.get("w")
        # cell centres: u[i][k][j] at array [i, k, j+1]; v[i][k][j] at [i, k+1, j]; w[i][k][j] at [i+1, k, j]
        uc = 0.5 * (u[1:n + 1, 1:n + 1, 2:n + 2] + u[1:n + 1, 1:n + 1, 1:n + 1])
        vc = 0.5 * (v[1:n + 1, 2:n + 2, 1:n + 1] + v[1:n + 1, 1:n + 1, 1:n + 1])
        wc = 0.5 * (w[2:n + 2, 1:n + 1, 1:n + 1] + w[1:n + 1, 1:n + 1, 1:n + 1])
        want = np.stack([uc, vc, wc], axis=-1)
        assert np.abs(vel - want).max() <= 1e-6, t     # %f prints 6 decimals
