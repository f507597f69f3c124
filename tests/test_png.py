"""PNG snapshots (fdm_amd/velocity_png.hpp) of the fdm_ns_cube driver.

CPU: the encoder, through tests/cpp/png_probe.cpp: a valid PNG (signature,
chunk CRCs, zlib stream, IHDR 784 x 524 RGB8), each panel's pixels the
palette colour of the 10-level bin of the known field value under them, the
titles in tEXt chunks.  The decoding here is Python's own zlib and struct.

GPU: the driver's snapshot after 10 steps of a 15^3 cavity against the CPU
oracle: the U, V (z = nz/2) and W (y = ny/2) panels and the three stream
function panels (psi from the oracle's LaplRect / LaplRectFFT2 solves of the
slice vorticity, as velocity_plotter::update, reference
src/velocity_plot.cpp:18-64), bin by bin; a pixel may differ only where the
value sits within 1e-9 of a bin edge.
"""
import os
import struct
import subprocess
import sys
import zlib

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROBE = os.path.join(ROOT, "tests", "cpp", "png_probe")
DRIVER = os.path.join(ROOT, "fdm_amd", "fdm_ns_cube")

PANEL, GAP, COLS, ROWS, LEVELS = 256, 4, 3, 2, 10
PALETTE = np.array([[49, 54, 149], [69, 117, 180], [116, 173, 209], [171, 217, 233], [224, 243, 248],
                    [254, 224, 144], [253, 174, 97], [244, 109, 67], [215, 48, 39], [165, 0, 38]], np.uint8)


def read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        typ, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert zlib.crc32(typ + body) & 0xffffffff == crc, typ
        chunks.append((typ, body))
        pos += 12 + n
    assert chunks[0][0] == b"IHDR" and chunks[-1][0] == b"IEND"
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, ctype, comp, filt, inter) == (8, 2, 0, 0, 0)
    raw = zlib.decompress(b"".join(b for t, b in chunks if t == b"IDAT"))
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 3 * w)
    assert (rows[:, 0] == 0).all()    # filter type none
    text = dict(b.split(b"\0", 1) for t, b in chunks if t == b"tEXt")
    return rows[:, 1:].reshape(h, w, 3), {k.decode(): v.decode() for k, v in text.items()}


def panel_levels(img, k):
    """level index per pixel of panel k (-1 for a colour not in the palette)."""
    x0 = GAP + (k % COLS) * (PANEL + GAP)
    y0 = GAP + (k // COLS) * (PANEL + GAP)
    px = img[y0:y0 + PANEL, x0:x0 + PANEL].reshape(-1, 1, 3)
    hit = (px == PALETTE[None]).all(-1)
    lev = np.where(hit.any(-1), hit.argmax(-1), -1)
    return lev.reshape(PANEL, PANEL)


def expected_levels(f):
    """the writer's rule: 10 equal bins over [min, max]; panel row 0 at the
    top shows the field's last row; pixel p shows cell p * n // 256."""
    lo, hi = float(f.min()), float(f.max())
    rows, cols = f.shape
    ri = (np.arange(PANEL)[::-1] * rows) // PANEL
    ci = (np.arange(PANEL) * cols) // PANEL
    v = f[np.ix_(ri, ci)]
    if not hi > lo:
        return np.zeros((PANEL, PANEL), int), v, lo, hi
    return np.clip(np.floor((v - lo) / (hi - lo) * LEVELS).astype(int), 0, LEVELS - 1), v, lo, hi


def check_panel(img, k, f, tol=1e-9):
    got = panel_levels(img, k)
    want, v, lo, hi = expected_levels(np.asarray(f, np.float64))
    assert (got >= 0).all(), k
    bad = got != want
    if bad.any():   # only values on a bin edge (to the tolerance) may land either side
        t = (v[bad] - lo) / (hi - lo) * LEVELS
        assert np.abs(t - np.round(t)).max() <= tol * LEVELS, (k, int(bad.sum()))


def test_png_encoder(tmp_path):
    if not os.path.exists(PROBE):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp"), PROBE])
    out = tmp_path / "p.png"
    subprocess.check_call([PROBE, str(out)], timeout=30)
    img, text = read_png(out)
    assert img.shape == (ROWS * PANEL + (ROWS + 1) * GAP, COLS * PANEL + (COLS + 1) * GAP, 3)
    assert text == {f"Panel{k}": f"title {k}" for k in range(6)}
    for k in range(6):
        rows, cols = 7 + k, 11 + 2 * k
        f = (np.arange(rows)[:, None] * cols + np.arange(cols)[None, :]) * (k + 1.0) - 3.0 * k
        check_panel(img, k, f, tol=0.0)
    assert (img[:GAP] == 255).all() and (img[:, :GAP] == 255).all()   # white gaps


def slices(u, v, w, nx, ny, nz):
    """velocity_plotter::update's cell-centred slices and vorticity RHS."""
    U = lambda i, k, j: u[i, k, j + 1]
    V = lambda i, k, j: v[i, k + 1, j]
    W = lambda i, k, j: w[i + 1, k, j]
    xm, ym, zm = nx // 2, ny // 2, nz // 2
    I, K = np.meshgrid(np.arange(nz + 2), np.arange(ny + 2), indexing="ij")
    vx = 0.5 * (V(I, K - 1, xm) + V(I, K, xm))
    wx = 0.5 * (W(I - 1, K, xm) + W(I, K, xm))
    I, J = np.meshgrid(np.arange(nz + 2), np.arange(nx + 2), indexing="ij")
    uy = 0.5 * (U(I, ym, J - 1) + U(I, ym, J))
    wy = 0.5 * (W(I - 1, ym, J) + W(I, ym, J))
    K, J = np.meshgrid(np.arange(ny + 2), np.arange(nx + 2), indexing="ij")
    uz = 0.5 * (U(zm, K, J - 1) + U(zm, K, J))
    vz = 0.5 * (V(zm, K - 1, J) + V(zm, K, J))
    return vx, wx, uy, wy, uz, vz


@pytest.mark.gpu
def test_driver_png_matches_oracle(tmp_path):
    import oracle as O
    n, steps = 15, 10
    print(f"[case] fdm_ns_cube PNG {n}^3 after {steps} steps", flush=True)
    sys.stdout.flush()
    if not os.path.exists(DRIVER):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "fdm_amd")])
    r = subprocess.run([DRIVER, f"--ns:nx={n}", f"--ns:nz={n}", "--ns:Re=250", "--ns:dt=0.01", f"--ns:steps={steps}",
                        "--plot:png=1", "--plot:vtk=0", f"--plot:interval={steps}"],
                       capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    assert (tmp_path / "step_0000000.png").exists()
    img, text = read_png(tmp_path / f"step_{steps:07d}.png")
    o = O.NsCube(n, n, n, u0=1.0, Re=250.0, dt=0.01)
    o.step(steps)
    u, v, w = o.get("u"), o.get("v"), o.get("w")
    nx = ny = nz = n
    vx, wx, uy, wy, uz, vz = slices(u, v, w, nx, ny, nz)
    h = 2 * np.pi / n
    rx = (wx[1:-1, 2:] - wx[1:-1, :-2]) / 2 / h - (vx[2:, 1:-1] - vx[:-2, 1:-1]) / 2 / h
    ry = (wy[1:-1, 2:] - wy[1:-1, :-2]) / 2 / h - (uy[2:, 1:-1] - uy[:-2, 1:-1]) / 2 / h
    rz = (vz[1:-1, 2:] - vz[1:-1, :-2]) / 2 / h - (uz[2:, 1:-1] - uz[:-2, 1:-1]) / 2 / h
    L = (n + 1) * h
    px = O.lapl_rect_fft2(rx, h, h, L, L)
    py = O.lapl_rect(ry, h, h, L, L)
    pz = O.lapl_rect(rz, h, h, L, L)
    for k, f in enumerate((uz, vz, wy, pz, py, px)):
        check_panel(img, k, f)
    t = steps * 0.01
    assert text["Panel0"] == "U (t=%.1e, |max|=%.1e)" % (t, np.abs(uz).max())
    assert text["Panel5"] == "VW (X=const) (%.1e)" % t
