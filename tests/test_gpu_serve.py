"""GPU tests of the batching / serving layer (fdmgpu_solve_table_device,
fdmgpu_group_*, fdmgpu_server_*, fdmgpu_tune, kernel variants).

Bars: every kernel variant meets the solver's parity bar against the oracle
(FP64 <= 1e-12 relative, FP32 <= 2e-6 vs the FP64 oracle, FP32+refine
residual <= 1e-10); a problem's answer is bitwise the same alone, in a
contiguous batch, through a pointer table, in a heterogeneous group (eager and
every graph replay) and among other in-flight server requests.
"""
import os
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def fdm():
    import fdm_amd
    return fdm_amd


@pytest.fixture(scope="module")
def hip():
    from fdm_amd import hip as h
    return h


@pytest.fixture(scope="module")
def srv():
    from fdm_amd import serve
    return serve


def rel2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def geom_cube(nx, ny, nz):
    h = 2 * np.pi / max(nx, ny, nz)
    return (h, h, h, (nx + 1) * h, (ny + 1) * h, (nz + 1) * h)


def geom_rect(nx, ny):
    h = 2 * np.pi / max(nx, ny)
    return (h, h, (nx + 1) * h, (ny + 1) * h)


def manufactured(shape, seed):
    """u seeded in [-1, 1), f = the discrete Laplacian of u (cube: 7-point, rect: 5-point)."""
    u = O.field(seed, shape)
    if len(shape) == 3:
        nz, ny, nx = shape
        h = geom_cube(nx, ny, nz)[0]
        f = O.apply_lapl(u, h, h, h)
    else:
        ny, nx = shape
        h = geom_rect(nx, ny)[0]
        f = O.apply_lapl(u, h, h)
    return u, f


# (255, 127) refine and (255, 31, 15) f32: FP32 y tiles of rows at least 128
# wide, 64 pencils per block by default
@pytest.mark.parametrize("case", [("cube", (31, 31, 31), "refine"), ("cube", (63, 31, 15), "f64"),
                                  ("cube", (63, 63, 63), "f32"), ("cube", (127, 127, 31), "refine"),
                                  ("rect", (100, 63), "refine"), ("rect", (255, 127), "f64"),
                                  ("rect", (255, 127), "refine"), ("cube", (255, 31, 15), "f32")])
def test_every_variant_parity(fdm, case):
    kind, shape, prec = case
    say("variants", case)
    if kind == "cube":
        nx, ny, nz = shape
        g = geom_cube(nx, ny, nz)
        s = fdm.LaplCube(*g, nx, ny, nz, precision=prec)
        arr = (nz, ny, nx)
        ref = O.lapl_cube(O.field(7, arr), *g)
    else:
        nx, ny = shape
        g = geom_rect(nx, ny)
        s = fdm.LaplRect(*g, nx, ny, precision=prec)
        arr = (ny, nx)
        ref = O.lapl_rect(O.field(7, arr), *g)
    rhs = O.field(7, arr)
    variants = s.variants()
    assert s.variant in variants
    for v in variants:
        s.variant = v
        say("  variant", v)
        if prec == "f32":
            r32 = rhs.astype(np.float32)
            ans = np.zeros_like(r32)
            s.solve(ans, r32)
            assert rel2(ans, ref) <= 2e-6, v
        else:
            ans = np.zeros_like(rhs)
            s.solve(ans, rhs)
            if prec == "f64":
                assert rel2(ans, ref) <= 1e-12, v
            else:
                st = s.stats()
                assert st.rel_residual <= 1e-10, v
                assert rel2(ans, ref) <= 1e-9, v
    with pytest.raises(fdm.FdmGpuError):
        s.variant = (0, 48, 48, 1)   # not a supported tile


def test_table_solve_matches_batch(fdm, hip):
    n, B = 31, 3
    say("table solve", n, B)
    g = geom_cube(n, n, n)
    F = np.stack([O.field(30 + i, (n, n, n)) for i in range(B)])
    s = fdm.LaplCube(*g, n, n, n, precision="refine")
    d_f = hip.DeviceArray.from_host(F)
    d_u = hip.DeviceArray(F.shape, np.float64)
    st = hip.Stream()
    s.solve_batch(d_u, d_f, st)
    st.synchronize()
    U = d_u.download()
    # the same problems in separate buffers, in reverse order, through a pointer table
    fs = [hip.DeviceArray.from_host(F[i]) for i in range(B)]
    us = [hip.DeviceArray((n, n, n), np.float64) for _ in range(B)]
    order = list(reversed(range(B)))
    tab_u = hip.DeviceArray.from_host(np.array([us[i].ptr.value for i in order], np.uint64))
    tab_f = hip.DeviceArray.from_host(np.array([fs[i].ptr.value for i in order], np.uint64))
    s.solve_table(tab_u, tab_f, B, st)
    st.synchronize()
    for i in range(B):
        assert np.array_equal(us[i].download(), U[i])


def _uneven_problems(srv, hip, fdm, count, seed, precs=("refine",)):
    """count problems: per axis n in {31, 63} (2-D rect for every 4th), FP32+refine
    (and FP64 when requested), manufactured rhs on the device."""
    rng = np.random.default_rng(seed)
    probs, keep = [], []
    for i in range(count):
        prec = precs[i % len(precs)]
        p = {"refine": fdm.F32_REFINE64, "f64": fdm.F64}[prec]
        if i % 4 == 3:
            nx, ny = (int(v) for v in rng.choice([31, 63], 2))
            u, f = manufactured((ny, nx), 100 + i)
            d_f = hip.DeviceArray.from_host(f)
            d_u = hip.DeviceArray(f.shape, np.float64)
            probs.append(srv.rect_problem(*geom_rect(nx, ny), nx, ny, d_u, d_f, precision=p))
        else:
            nx, ny, nz = (int(v) for v in rng.choice([15, 31, 63], 3))
            u, f = manufactured((nz, ny, nx), 100 + i)
            d_f = hip.DeviceArray.from_host(f)
            d_u = hip.DeviceArray(f.shape, np.float64)
            probs.append(srv.cube_problem(*geom_cube(nx, ny, nz), nx, ny, nz, d_u, d_f, precision=p))
        keep.append((d_u, d_f, u, f, prec))
    return probs, keep


def test_group_uneven_eager_and_graph_replays(fdm, hip, srv):
    say("group uneven")
    probs, keep = _uneven_problems(srv, hip, fdm, 24, seed=3, precs=("refine", "refine", "f64"))
    grp = srv.Group(probs, nstreams=4, graph=True)
    st = hip.Stream()
    grp.solve(st)
    st.synchronize()
    eager = [k[0].download() for k in keep]
    for i, (d_u, d_f, u, f, prec) in enumerate(keep):
        r = grp.result(i)
        assert r.status == 0
        if prec == "refine":
            assert r.rel_residual <= 1e-10
        # manufactured: the solve recovers u (discrete problem, exact up to rounding)
        assert rel2(eager[i], u) <= (1e-9 if prec == "refine" else 1e-12), (i, prec)
    for rep in range(3):          # graph capture on the second solve, replays after
        for d_u, *_ in keep:      # scribble over the answers (eager work between replays)
            d_u.upload(np.full(d_u.shape, 7.0))
        hip.synchronize()
        grp.solve(st)
        st.synchronize()
        for i, (d_u, *_rest) in enumerate(keep):
            assert np.array_equal(d_u.download(), eager[i]), (rep, i)
    grp.close()


@pytest.mark.parametrize("nstreams", [4, 2, 1])
def test_server_alone_vs_mix(fdm, hip, srv, nstreams):
    """Results are bitwise the same alone and in the mix, with the default
    four streams (the first, for large requests, on a CU mask), two, and one
    stream shared by every size class (no mask)."""
    say("server alone vs mix", nstreams)
    probs, keep = _uneven_problems(srv, hip, fdm, 12, seed=5)
    noise, nkeep = _uneven_problems(srv, hip, fdm, 12, seed=6)
    S = srv.Server(hbm_budget=512 << 20, max_batch=8, nstreams=nstreams)
    alone = []
    for p, k in zip(probs, keep):
        t = S.submit(p)
        r = S.wait(t)
        assert r.state == 2 and r.status == 0 and r.rel_residual <= 1e-10
        alone.append(k[0].download())
    for d_u, *_ in keep:
        d_u.upload(np.zeros(d_u.shape))
    hip.synchronize()
    tickets = []
    for p, q in zip(probs, noise):    # interleaved with other requests, all in flight together
        tickets.append(S.submit(p))
        S.submit(q)
    S.drain()
    for i, t in enumerate(tickets):
        r = S.poll(t)
        assert r.state == 2 and r.status == 0
        assert r.latency_ms >= r.queue_ms >= 0
        assert np.array_equal(keep[i][0].download(), alone[i]), i
        assert rel2(alone[i], keep[i][2]) <= 1e-9
    S.close()


def test_server_admission(fdm, hip, srv):
    say("server admission")
    n = 63
    u, f = manufactured((n, n, n), 9)
    d_f = hip.DeviceArray.from_host(f)
    outs = [hip.DeviceArray(f.shape, np.float64) for _ in range(3)]
    g = geom_cube(n, n, n)
    small = srv.Server(hbm_budget=64 << 10, max_batch=4)
    with pytest.raises(fdm.FdmGpuError) as e:      # scratch of one request exceeds the budget
        small.submit(srv.cube_problem(*g, n, n, n, outs[0], d_f))
    assert e.value.code == fdm.ERR_ADMISSION
    small.close()
    # a budget for one request's scratch at a time: three requests still complete, one by one
    resident, per_solve = fdm.LaplCube(*g, n, n, n, precision="refine").footprint(1)
    assert per_solve >= 2 * n ** 3 * 4        # FP32 rhs/correction and first iterate
    one = per_solve + (256 << 10)
    S = srv.Server(hbm_budget=one, max_batch=4)
    ts = [S.submit(srv.cube_problem(*g, n, n, n, o, d_f)) for o in outs]
    S.drain()
    for t, o in zip(ts, outs):
        r = S.poll(t)
        assert r.status == 0 and r.rel_residual <= 1e-10
        assert rel2(o.download(), u) <= 1e-9
    S.close()


def test_tune_writes_cache_and_applies(fdm, hip, srv, tmp_path):
    say("tune")
    probs, keep = _uneven_problems(srv, hip, fdm, 4, seed=8)
    path = str(tmp_path / "tune.json")
    srv.tune(probs, reps=2, cache_path=path)
    text = open(path).read()
    assert text.count('"plane"') >= 1
    # a new plan of a tuned shape starts from the tuned variant
    p = probs[0]
    if p.kind == srv.KIND_CUBE:
        s = fdm.LaplCube(p.dx, p.dy, p.dz, p.lx, p.ly, p.lz, p.nx, p.ny, p.nz, precision="refine")
    else:
        s = fdm.LaplRect(p.dx, p.dy, p.lx, p.ly, p.nx, p.ny, precision="refine")
    kind = "cube" if p.kind == srv.KIND_CUBE else "rect"
    key = f'"{kind}:p2:m0:{p.nx}x{p.ny}x{p.nz if p.kind == srv.KIND_CUBE else 1}"'
    line = [l for l in text.splitlines() if l.strip().startswith(key)][0]
    v = s.variant
    assert f'"plane": {v[0]}, "tile_x": {v[1]}, "tile_y": {v[2]}, "zcr_gmul": {v[3]}' in line


@pytest.mark.parametrize("shape,fuse", [((31, 31, 31), 0), ((31, 31, 31), 1), ((63, 31, 15), 0)])
def test_server_continuation_matches_uninterrupted(fdm, hip, srv, shape, fuse):
    """A request that needs more inner solves than its solver's earlier requests
    continues from its iterate in a follow-up batch; the answer is bitwise the
    uninterrupted solve's (plane path with the separate and the fused
    residual, and the tile path)."""
    nx, ny, nz = shape
    say("server continuation", shape, fuse)
    g = geom_cube(nx, ny, nz)
    u, f = manufactured((nz, ny, nx), 41)
    s = fdm.LaplCube(*g, nx, ny, nz, precision="refine")
    d_x, d_y = hip.DeviceArray(f.shape, np.float64), hip.DeviceArray(f.shape, np.float64)
    shape_problem = srv.cube_problem(*g, nx, ny, nz, d_x, d_y)
    default = s.variant
    if nx == ny:   # pin the residual variant for this shape (plans and server batches)
        v = next(x for x in s.variants() if x[0] == 1 and x[4] == fuse)
        srv.tune_set(shape_problem, v)
        s.variant = v
    try:
        _continuation_case(fdm, hip, srv, s, g, nx, ny, nz, f)
    finally:
        if nx == ny:
            srv.tune_set(shape_problem, default)


def _continuation_case(fdm, hip, srv, s, g, nx, ny, nz, f):
    s.set_refine(1e-30, 2)
    ans = np.zeros_like(f)
    with pytest.raises(fdm.FdmGpuError):        # 2 inner solves cannot reach 1e-30
        s.solve(ans, f)
    r2 = s.stats().rel_residual
    tol = r2 / 2                                # 2 solves miss it, 3 reach it
    s.set_refine(tol, 4)
    ref = np.zeros_like(f)
    s.solve(ref, f)
    assert s.stats().iterations == 3
    d_f = hip.DeviceArray.from_host(f)
    d_a, d_b = hip.DeviceArray(f.shape, np.float64), hip.DeviceArray(f.shape, np.float64)
    S = srv.Server(hbm_budget=256 << 20, max_batch=4)
    r = S.wait(S.submit(srv.cube_problem(*g, nx, ny, nz, d_a, d_f)))   # the solver needs 2 at 1e-10
    assert r.status == 0 and r.iterations == 2
    S.set_refine(tol, 4)
    r = S.wait(S.submit(srv.cube_problem(*g, nx, ny, nz, d_b, d_f)))   # launched with 2, continued to 3
    assert r.status == 0 and r.iterations == 3 and r.rel_residual <= tol
    assert np.array_equal(d_b.download(), ref)
    S.close()


def test_tune_set_rejects_invalid_variant(fdm, hip, srv):
    """fdmgpu_tune_set accepts only variants valid for the shape."""
    say("tune_set invalid")
    g = geom_cube(255, 31, 15)
    d_x, d_y = hip.DeviceArray((15, 31, 255), np.float64), hip.DeviceArray((15, 31, 255), np.float64)
    p = srv.cube_problem(*g, 255, 31, 15, d_x, d_y)
    with pytest.raises(fdm.FdmGpuError) as e:      # 255-wide rows: no plane kernel (axis tiles only)
        srv.tune_set(p, (1, 0, 0, 0, 0))
    assert e.value.code == fdm.ERR_INVALID_ARG
    with pytest.raises(fdm.FdmGpuError):           # no such tile
        srv.tune_set(p, (0, 48, 0, 0, 0))
