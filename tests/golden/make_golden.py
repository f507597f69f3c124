"""Write tests/golden/ref_vectors.npz from the REFERENCE's own code.

Runs in the container that holds the reference checkout: it loads
oracle/_ref/libfdmref.so (built by `make -C oracle` from the reference's own
sources, see oracle/Makefile) and records seeded inputs and the reference's
outputs as small fixtures.  The fixtures are data (inputs + expected outputs),
not reference source; the tests compare the oracle and the GPU library with
them on machines where the reference is absent.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402


def main():
    R = O.ref()
    out = {}
    rng = np.random.default_rng(20261202)

    # 1-D transforms (src/fft.cpp): kind 0 sFFT, 1 cFFT, 2 pFFT_1, 3 pFFT
    for N in (8, 32, 128):
        for dt, tag, fn in ((np.float64, "d", R.ref_fft_d), (np.float32, "f", R.ref_fft_f)):
            x = rng.uniform(-1, 1, N + 1).astype(dt)
            for kind, name in ((0, "sfft"), (1, "cfft"), (2, "pfft1"), (3, "pfft")):
                S = np.zeros(N + 1, dt)
                fn(S, x.copy(), N, dt(0.37), kind)
                out[f"{name}_{tag}_N{N}_in"] = x
                out[f"{name}_{tag}_N{N}_out"] = S

    # LaplCube::solve (src/lapl_cube.cpp:9-142), Dirichlet, NSCube geometry
    for n in (15, 31):
        u, f, h, l = O.manufactured_cube(n, seed=1)
        for dt, tag in ((np.float64, "d"), (np.float32, "f")):
            rhs = f.astype(dt)
            ans = O.ref_lapl_cube(rhs, h, h, h, l, l, l)
            out[f"cube_{tag}_{n}_rhs"] = rhs
            out[f"cube_{tag}_{n}_ans"] = ans
            out[f"cube_{tag}_{n}_geom"] = np.array([h, h, h, l, l, l])

    # anisotropic box with the eigenvalue-aliasing quirk (nx == ny, dx != dy)
    nx, ny, nz = 15, 15, 7
    dx, dy, dz = 0.11, 0.07, 0.13
    lx, ly, lz = (nx + 1) * dx, (ny + 1) * dy, (nz + 1) * dz
    rhs = O.field(7, (nz, ny, nx))
    out["cube_aniso_d_rhs"] = rhs
    out["cube_aniso_d_ans"] = O.ref_lapl_cube(rhs, dx, dy, dz, lx, ly, lz)
    out["cube_aniso_d_geom"] = np.array([dx, dy, dz, lx, ly, lz])

    # periodic LaplCube (n-body path, test/nbody.cpp:323)
    for N in (16,):
        h = 2 * np.pi / N
        g = O.field(5, (N, N, N))
        for dt, tag in ((np.float64, "d"), (np.float32, "f")):
            rhs = g.astype(dt)
            out[f"cubep_{tag}_{N}_rhs"] = rhs
            out[f"cubep_{tag}_{N}_ans"] = O.ref_lapl_cube(rhs, h, h, h, 2 * np.pi, 2 * np.pi, 2 * np.pi,
                                                          periodic=True)
            out[f"cubep_{tag}_{N}_geom"] = np.array([h, h, h, 2 * np.pi, 2 * np.pi, 2 * np.pi])

    # tridiagonal solvers (src/asp_gauss.cpp:185-229, src/cyclic_reduction.h)
    for n, general in ((31, 0), (127, 0), (20, 1), (100, 1)):
        A1 = rng.uniform(-1, 1, n)
        A2 = rng.uniform(-1, 1, n) + np.sign(rng.uniform(-1, 1, n)) * 3.0   # diagonally dominant
        A3 = rng.uniform(-1, 1, n)
        B = rng.uniform(-1, 1, n)
        e = np.concatenate([[0.0], A1[:-1]])   # CR lower: e[0] = 0, e[j] couples x_{j-1}
        x_thomas = B.copy()
        R.ref_thomas_d(x_thomas, A1.copy(), A2.copy(), A3.copy(), n)
        x_cr = B.copy()
        q = int(np.ceil(np.log2(n + 1)))
        R.ref_cr_d(A2.copy(), e.copy(), A3.copy(), x_cr, q, n, general)
        x_cls = B.copy()
        R.ref_cr_class_d(A2.copy(), e.copy(), A3.copy(), x_cls, n)
        key = f"tdiag_{n}"
        out[key + "_sub"] = A1
        out[key + "_diag"] = A2
        out[key + "_sup"] = A3
        out[key + "_rhs"] = B
        out[key + "_thomas"] = x_thomas
        out[key + "_cr"] = x_cr
        out[key + "_crclass"] = x_cls
        out[key + "_general"] = np.array([general])

    path = os.path.join(HERE, "ref_vectors.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {len(out)} arrays, {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
