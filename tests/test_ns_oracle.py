"""CPU checks of the oracle's NSCube step (restatement of src/ns_cube.cpp;
the reference's ns_cube.cpp needs plplot and cannot be built here).  Sizes
have n + 1 a power of two (the Poisson solve's transform length).

Pins: a closed form of the first step from rest, rest staying at rest,
discrete incompressibility of the interior after every projection, closed
forms of FGH (src/ns_cube.cpp:126-200) -- affine velocity fields, for which
every averaged product term is exact (all nine advective terms), and fields
quadratic in one coordinate (all nine viscous terms) -- and an independent
array-slicing transcription of init_bound / FGH / the Poisson RHS /
update_uvwp (src/ns_cube.cpp:64-277) compared with the oracle's loops.  The
FGH closed forms fail when an index, a spacing or a term of the oracle's FGH
is changed (checked by planting such changes in a copy).
"""
import numpy as np
import pytest

import oracle as O


def interior_divergence(ns):
    u, v, w = ns.get("u"), ns.get("v"), ns.get("w")
    nx, ny, nz = ns.nx, ns.ny, ns.nz
    # u[i][k][j] at array [i, k, j+1]; v[i][k][j] at [i, k+1, j]; w[i][k][j] at [i+1, k, j]
    div = ((u[1:nz + 1, 1:ny + 1, 2:nx + 2] - u[1:nz + 1, 1:ny + 1, 1:nx + 1]) / ns.dx
           + (v[1:nz + 1, 2:ny + 2, 1:nx + 1] - v[1:nz + 1, 1:ny + 1, 1:nx + 1]) / ns.dy
           + (w[2:nz + 2, 1:ny + 1, 1:nx + 1] - w[1:nz + 1, 1:ny + 1, 1:nx + 1]) / ns.dz)
    return div


def test_ns_first_step_closed_form():
    """From rest only the lid acts: after one step u = 2 U0 dt / (Re dz^2) on the
    top layer i = nz (faces j = 1..nx-1), everything else zero."""
    n, Re, dt, u0 = 15, 250.0, 0.01, 1.0
    ns = O.NsCube(n, n, n, u0=u0, Re=Re, dt=dt)
    ns.step(1)
    u = ns.get("u")
    c = 2 * u0 * dt / (Re * ns.dz ** 2)
    want = np.zeros_like(u)
    want[n, 1:n + 1, 2:n + 1] = c                 # i = nz, k = 1..ny, j = 1..nx-1
    want[n + 1, :, :] = 2 * u0                    # lid ghost row: 2 U0 - u(nz) with u(nz) = 0 before the step
    want[n + 1, :, 0] = want[n + 1, :, 2]         # x ghosts of the lid row
    want[n + 1, :, n + 2] = want[n + 1, :, n]
    assert np.allclose(u[1:n + 1], want[1:n + 1], rtol=0, atol=1e-15 * c)
    assert np.array_equal(u[n + 1], want[n + 1])
    for f in ("v", "w", "p", "x"):
        assert np.abs(ns.get(f)).max() == 0.0, f


def test_ns_zero_lid_stays_at_rest():
    ns = O.NsCube(15, 15, 7, u0=0.0, Re=100.0, dt=0.01)
    ns.step(5)
    for f in ("u", "v", "w", "p"):
        assert np.abs(ns.get(f)).max() == 0.0, f


@pytest.mark.parametrize("shape", [(15, 15, 15), (31, 15, 7)])
def test_ns_interior_divergence_free(shape):
    nx, ny, nz = shape
    ns = O.NsCube(nx, ny, nz, u0=1.0, Re=250.0, dt=0.01)
    for _ in range(5):
        ns.step(2)
        d = interior_divergence(ns)
        scale = np.abs(ns.get("u")[1:nz + 1]).max() / ns.dx
        # cells whose six faces are all updated by the projection
        assert np.abs(d[1:-1, 1:-1, 1:-1]).max() <= 1e-12 * scale


# ---- independent transcription with array slices ---------------------------
def _np_step(st, p_, dx, dy, dz, dt, Re, u0):
    """One NSCube step on padded arrays: u[i, k, j+1], v[i, k+1, j], w[i+1, k, j],
    p[i, k, j] (reference index ranges, src/ns_cube.h:66-75)."""
    u, v, w, p = st["u"].copy(), st["v"].copy(), st["w"].copy(), st["p"].copy()
    nz, ny, nx = p.shape[0] - 2, p.shape[1] - 2, p.shape[2] - 2
    # init_bound
    u[nz + 1, :, 0:nz + 3] = 2 * u0 - u[nz, :, 0:nz + 3]
    u[:, :, 0] = u[:, :, 2]
    u[:, :, nx + 2] = u[:, :, nx]
    v[:, 0, :] = v[:, 2, :]
    v[:, ny + 2, :] = v[:, ny, :]
    w[0, :, :] = w[2, :, :]
    w[nz + 2, :, :] = w[nz, :, :]
    I, K, J = slice(1, nz + 1), slice(1, ny + 1), slice(1, nx + 1)
    p[I, K, 0] = p[I, K, 1] - (u[I, K, 2] - 2 * u[I, K, 1] + u[I, K, 0]) / Re / dx
    p[I, K, nx + 1] = p[I, K, nx] - (u[I, K, nx + 2] - 2 * u[I, K, nx + 1] + u[I, K, nx]) / Re / dx
    p[I, 0, J] = p[I, 1, J] - (v[I, 2, J] - 2 * v[I, 1, J] + v[I, 0, J]) / Re / dy
    p[I, ny + 1, J] = p[I, ny, J] - (v[I, ny + 2, J] - 2 * v[I, ny + 1, J] + v[I, ny, J]) / Re / dy
    p[0, K, J] = p[1, K, J] - (w[2, K, J] - 2 * w[1, K, J] + w[0, K, J]) / Re / dz
    p[nz + 1, K, J] = p[nz, K, J] - (w[nz + 2, K, J] - 2 * w[nz + 1, K, J] + w[nz, K, J]) / Re / dz

    def U(di, dk, dj, ir, kr, jr):     # u at (i+di, k+dk, j+dj) over index ranges (lo, hi) inclusive
        return u[ir[0] + di:ir[1] + di + 1, kr[0] + dk:kr[1] + dk + 1, jr[0] + dj + 1:jr[1] + dj + 2]

    def Vv(di, dk, dj, ir, kr, jr):
        return v[ir[0] + di:ir[1] + di + 1, kr[0] + dk + 1:kr[1] + dk + 2, jr[0] + dj:jr[1] + dj + 1]

    def Ww(di, dk, dj, ir, kr, jr):
        return w[ir[0] + di + 1:ir[1] + di + 2, kr[0] + dk:kr[1] + dk + 1, jr[0] + dj:jr[1] + dj + 1]

    dx2, dy2, dz2 = dx * dx, dy * dy, dz * dz
    r = ((1, nz), (1, ny), (0, nx))
    uc = U(0, 0, 0, *r)
    F = uc + dt * ((U(0, 0, 1, *r) - 2 * uc + U(0, 0, -1, *r)) / Re / dx2
                   + (U(0, 1, 0, *r) - 2 * uc + U(0, -1, 0, *r)) / Re / dy2
                   + (U(1, 0, 0, *r) - 2 * uc + U(-1, 0, 0, *r)) / Re / dz2
                   - ((0.5 * (uc + U(0, 0, 1, *r))) ** 2 - (0.5 * (U(0, 0, -1, *r) + uc)) ** 2) / dx
                   - 0.25 * ((uc + U(0, 1, 0, *r)) * (Vv(0, 0, 1, *r) + Vv(0, 0, 0, *r))
                             - (U(0, -1, 0, *r) + uc) * (Vv(0, -1, 1, *r) + Vv(0, -1, 0, *r))) / dy
                   - 0.25 * ((uc + U(1, 0, 0, *r)) * (Ww(0, 0, 1, *r) + Ww(0, 0, 0, *r))
                             - (U(-1, 0, 0, *r) + uc) * (Ww(-1, 0, 1, *r) + Ww(-1, 0, 0, *r))) / dz)
    r = ((1, nz), (0, ny), (1, nx))
    vc = Vv(0, 0, 0, *r)
    G = vc + dt * ((Vv(0, 0, 1, *r) - 2 * vc + Vv(0, 0, -1, *r)) / Re / dx2
                   + (Vv(0, 1, 0, *r) - 2 * vc + Vv(0, -1, 0, *r)) / Re / dy2
                   + (Vv(1, 0, 0, *r) - 2 * vc + Vv(-1, 0, 0, *r)) / Re / dz2
                   - ((0.5 * (vc + Vv(0, 1, 0, *r))) ** 2 - (0.5 * (Vv(0, -1, 0, *r) + vc)) ** 2) / dy
                   - 0.25 * ((U(0, 0, 0, *r) + U(0, 1, 0, *r)) * (Vv(0, 0, 1, *r) + vc)
                             - (U(0, 0, -1, *r) + U(0, 1, -1, *r)) * (vc + Vv(0, 0, -1, *r))) / dx
                   - 0.25 * ((Ww(0, 0, 0, *r) + Ww(0, 1, 0, *r)) * (vc + Vv(1, 0, 0, *r))
                             - (Ww(-1, 0, 0, *r) + Ww(-1, 1, 0, *r)) * (Vv(-1, 0, 0, *r) + vc)) / dz)
    r = ((0, nz), (1, ny), (1, nx))
    wc = Ww(0, 0, 0, *r)
    H = wc + dt * ((Ww(0, 0, 1, *r) - 2 * wc + Ww(0, 0, -1, *r)) / Re / dx2
                   + (Ww(0, 1, 0, *r) - 2 * wc + Ww(0, -1, 0, *r)) / Re / dy2
                   + (Ww(1, 0, 0, *r) - 2 * wc + Ww(-1, 0, 0, *r)) / Re / dz2
                   - ((0.5 * (Ww(1, 0, 0, *r) + wc)) ** 2 - (0.5 * (Ww(-1, 0, 0, *r) + wc)) ** 2) / dz
                   - 0.25 * ((U(1, 0, 0, *r) + U(0, 0, 0, *r)) * (Ww(0, 0, 1, *r) + wc)
                             - (U(1, 0, -1, *r) + U(0, 0, -1, *r)) * (wc + Ww(0, 0, -1, *r))) / dx
                   - 0.25 * ((wc + Ww(0, 1, 0, *r)) * (Vv(0, 0, 0, *r) + Vv(1, 0, 0, *r))
                             - (Ww(0, -1, 0, *r) + wc) * (Vv(0, -1, 0, *r) + Vv(1, -1, 0, *r))) / dy)
    # RHS with the pressure boundary values moved to the right-hand side
    rhs = ((F[:, :, 1:] - F[:, :, :-1]) / dx + (G[:, 1:, :] - G[:, :-1, :]) / dy
           + (H[1:, :, :] - H[:-1, :, :]) / dz) / dt
    rhs[0] -= p[0, K, J] / dz2
    rhs[:, 0] -= p[I, 0, J] / dy2
    rhs[:, :, 0] -= p[I, K, 0] / dx2
    rhs[:, :, -1] -= p[I, K, nx + 1] / dx2
    rhs[:, -1] -= p[I, ny + 1, J] / dy2
    rhs[-1] -= p[nz + 1, K, J] / dz2
    lx, ly, lz = (nx + 1) * dx, (ny + 1) * dy, (nz + 1) * dz
    x = O.lapl_cube(rhs, dx, dy, dz, lx, ly, lz)
    u[1:nz + 1, 1:ny + 1, 2:nx + 1] = F[:, :, 1:nx] - dt / dx * (x[:, :, 1:] - x[:, :, :-1])
    v[1:nz + 1, 2:ny + 1, 1:nx + 1] = G[:, 1:ny, :] - dt / dy * (x[:, 1:, :] - x[:, :-1, :])
    w[2:nz + 1, 1:ny + 1, 1:nx + 1] = H[1:nz, :, :] - dt / dz * (x[1:, :, :] - x[:-1, :, :])
    p[I, K, J] = x
    return {"u": u, "v": v, "w": w, "p": p, "F": F, "G": G, "H": H, "RHS": rhs, "x": x}


def test_ns_step_vs_slicing_transcription():
    nx, ny, nz = 15, 7, 7
    Re, dt, u0 = 100.0, 0.005, 0.7
    ns = O.NsCube(nx, ny, nz, u0=u0, Re=Re, dt=dt)
    st = {}
    for i, f in enumerate(("u", "v", "w", "p")):      # an arbitrary state, not only the cavity's
        a = 0.1 * O.field(40 + i, ns.shapes[f])
        ns.set(f, a)
        st[f] = a
    ns.step(1)
    ref = _np_step(st, None, ns.dx, ns.dy, ns.dz, dt, Re, u0)
    for f in ("F", "G", "H", "RHS", "x", "u", "v", "w", "p"):
        got, want = ns.get(f), ref[f]
        assert np.allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max()), f


# ---------------------------------------------------------------------------
# FGH (src/ns_cube.cpp:126-200) pinned by closed forms.  On the staggered grid
# u[i][k][j] sits at (x, y, z) = (j dx, (k - 1/2) dy, (i - 1/2) dz), v[i][k][j]
# at ((j - 1/2) dx, k dy, (i - 1/2) dz), w[i][k][j] at ((j - 1/2) dx,
# (k - 1/2) dy, i dz).  For velocity fields affine in (x, y, z) the two-point
# averages of FGH are the fields at the midpoints and every difference of a
# product is a central difference of a quadratic, so the discrete update is
# exact: F = u + dt (-d(u^2)/dx - d(uv)/dy - d(uw)/dz), and likewise G, H (the
# viscous terms vanish).  For a field quadratic in one coordinate and the
# other components zero, only the viscous term along that coordinate is left:
# F = u + dt/Re d2u/dy2 for u = a y^2, etc.  Only points whose stencils stay
# clear of the boundary rows that init_bound rewrites are compared.
# ---------------------------------------------------------------------------
NS_GEOM = dict(x1=-np.pi, x2=np.pi, y1=-2.0, y2=2.0, z1=-1.0, z2=1.5)   # dx, dy, dz all different


def staggered_coords(ns, which):
    """Physical coordinates (z, y, x arrays) of every entry of u, v or w in its
    array layout (u: j from -1, v: k from -1, w: i from -1)."""
    nx, ny, nz = ns.nx, ns.ny, ns.nz
    shp = ns.shapes[which]
    ia, ka, ja = np.meshgrid(np.arange(shp[0]), np.arange(shp[1]), np.arange(shp[2]), indexing="ij")
    i, k, j = ia.astype(float), ka.astype(float), ja.astype(float)
    if which == "u":
        j = j - 1
        return (i - 0.5) * ns.dz, (k - 0.5) * ns.dy, j * ns.dx
    if which == "v":
        k = k - 1
        return (i - 0.5) * ns.dz, k * ns.dy, (j - 0.5) * ns.dx
    i = i - 1
    return i * ns.dz, (k - 0.5) * ns.dy, (j - 0.5) * ns.dx


def run_fgh(ns, fields):
    for name, f in fields.items():
        Z, Y, X = staggered_coords(ns, name)
        ns.set(name, f(X, Y, Z))
    ns.step(1)
    return ns.get("F"), ns.get("G"), ns.get("H")


def deep(a, lo=2, hi=2):
    return a[lo:a.shape[0] - hi, lo:a.shape[1] - hi, lo:a.shape[2] - hi]


def test_ns_fgh_affine_advection_closed_form():
    Re, dt = 50.0, 0.02
    ns = O.NsCube(15, 15, 7, u0=0.0, Re=Re, dt=dt, **NS_GEOM)
    a = (0.3, 0.7, -0.4, 0.9)     # u = a0 + ax x + ay y + az z
    b = (-0.2, 0.5, 0.8, -0.6)    # v
    c = (0.1, -0.9, 0.35, 0.45)   # w

    def aff(q):
        return lambda X, Y, Z: q[0] + q[1] * X + q[2] * Y + q[3] * Z

    F, G, H = run_fgh(ns, {"u": aff(a), "v": aff(b), "w": aff(c)})
    nx, ny, nz = ns.nx, ns.ny, ns.nz
    # F[i][k][j] (i 1..nz, k 1..ny, j 0..nx) at array [i-1, k-1, j], at the u point
    i, k, j = np.meshgrid(np.arange(1, nz + 1), np.arange(1, ny + 1), np.arange(0, nx + 1), indexing="ij")
    X, Y, Z = j * ns.dx, (k - 0.5) * ns.dy, (i - 0.5) * ns.dz
    u, v, w = aff(a)(X, Y, Z), aff(b)(X, Y, Z), aff(c)(X, Y, Z)
    wantF = u + dt * (-2 * u * a[1] - (a[2] * v + u * b[2]) - (a[3] * w + u * c[3]))
    # G[i][k][j] (i 1..nz, k 0..ny, j 1..nx) at [i-1, k, j-1], at the v point
    i, k, j = np.meshgrid(np.arange(1, nz + 1), np.arange(0, ny + 1), np.arange(1, nx + 1), indexing="ij")
    X, Y, Z = (j - 0.5) * ns.dx, k * ns.dy, (i - 0.5) * ns.dz
    u, v, w = aff(a)(X, Y, Z), aff(b)(X, Y, Z), aff(c)(X, Y, Z)
    wantG = v + dt * (-2 * v * b[2] - (a[1] * v + u * b[1]) - (b[3] * w + v * c[3]))
    # H[i][k][j] (i 0..nz, k 1..ny, j 1..nx) at [i, k-1, j-1], at the w point
    i, k, j = np.meshgrid(np.arange(0, nz + 1), np.arange(1, ny + 1), np.arange(1, nx + 1), indexing="ij")
    X, Y, Z = (j - 0.5) * ns.dx, (k - 0.5) * ns.dy, i * ns.dz
    u, v, w = aff(a)(X, Y, Z), aff(b)(X, Y, Z), aff(c)(X, Y, Z)
    wantH = w + dt * (-2 * w * c[3] - (a[1] * w + u * c[1]) - (b[2] * w + v * c[2]))
    for got, want in ((F, wantF), (G, wantG), (H, wantH)):
        assert got.shape == want.shape
        scale = np.abs(want).max()
        assert np.abs(deep(got) - deep(want)).max() <= 1e-13 * scale


@pytest.mark.parametrize("comp", ["u", "v", "w"])
@pytest.mark.parametrize("axis", ["x", "y", "z"])
def test_ns_fgh_viscous_closed_form(comp, axis):
    """One component q s^2 (s = x, y or z), the others 0.  The advective terms
    are even in q and the viscous ones odd, so half the difference of the runs
    with q and -q is u + dt/Re d2u/ds2 = q s^2 + 2 q dt / Re, for every one of
    the nine (component, axis) viscous terms."""
    Re, dt, q = 7.0, 0.01, 0.8
    got = []
    for sgn in (1.0, -1.0):
        ns = O.NsCube(15, 15, 7, u0=0.0, Re=Re, dt=dt, **NS_GEOM)
        fields = {"u": lambda X, Y, Z: 0 * X, "v": lambda X, Y, Z: 0 * X, "w": lambda X, Y, Z: 0 * X}
        ax = {"x": 0, "y": 1, "z": 2}[axis]
        fields[comp] = lambda X, Y, Z, ax=ax, sgn=sgn: sgn * q * (X, Y, Z)[ax] ** 2
        F, G, H = run_fgh(ns, fields)
        got.append({"u": F, "v": G, "w": H}[comp])
        for other in {"u": (G, H), "v": (F, H), "w": (F, G)}[comp]:
            assert np.abs(deep(other)).max() == 0.0
    odd = 0.5 * (got[0] - got[1])
    nx, ny, nz = ns.nx, ns.ny, ns.nz
    if comp == "u":
        i, k, j = np.meshgrid(np.arange(1, nz + 1), np.arange(1, ny + 1), np.arange(0, nx + 1), indexing="ij")
        X, Y, Z = j * ns.dx, (k - 0.5) * ns.dy, (i - 0.5) * ns.dz
    elif comp == "v":
        i, k, j = np.meshgrid(np.arange(1, nz + 1), np.arange(0, ny + 1), np.arange(1, nx + 1), indexing="ij")
        X, Y, Z = (j - 0.5) * ns.dx, k * ns.dy, (i - 0.5) * ns.dz
    else:
        i, k, j = np.meshgrid(np.arange(0, nz + 1), np.arange(1, ny + 1), np.arange(1, nx + 1), indexing="ij")
        X, Y, Z = (j - 0.5) * ns.dx, (k - 0.5) * ns.dy, i * ns.dz
    want = q * (X, Y, Z)[{"x": 0, "y": 1, "z": 2}[axis]] ** 2 + dt * 2 * q / Re
    assert np.abs(deep(odd) - deep(want)).max() <= 1e-12 * np.abs(want).max()
