// C ABI of the MI355X fast Poisson solve (include/fdmgpu.h): plans, solves,
// FP32+FP64-refinement driver.
//
// Solve pipeline for a Dirichlet cube (reference LaplCube::solve,
// src/lapl_cube.cpp:9-142):
//   1. DST-I along x and y of every z-plane, factors dx*slx, dy*sly
//      (reference passes 3 and 2)                 -> plane kernel, or x + y tiles
//   2. per (kx,ky) mode, the z tridiagonal system  -> zcr kernel
//      (replaces reference passes 1, 4 and 7: z DST, divide by -sum(lambda),
//      inverse z DST; same linear system, see dst_device.hpp)
//   3. inverse DST-I along x and y, factors slx, sly (reference passes 5, 6)
// The solve runs in place in `ans` (no scratch) for plain F32/F64 plans.
// LaplRect (src/lapl_rect.cpp:63-110): y DST -> per-row cyclic reduction of
// the gtsv system -> inverse y DST.
// Periodic axes (tensor_flags<periodic...>) use the half-complex real DFT
// pair (pFFT_1 / pFFT) instead of the DST, and the periodic cube's z systems
// are circulant (see zcrp_kernel).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "kernels.hpp"
#include "plan.hpp"
#include "refine.hpp"

using namespace fdmgpu;

namespace {

thread_local char g_errbuf[256];

int ilog2_exact(int N) {
    if (N < 1 || (N & (N - 1))) return -1;
    int l = 0;
    while ((1 << l) < N) l++;
    return l;
}

int hip_fail(hipError_t e) {
    std::snprintf(g_errbuf, sizeof g_errbuf, "HIP error %d: %s", (int)e, hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? FDMGPU_ERR_OUT_OF_MEMORY : FDMGPU_ERR_DEVICE;
}

#define HIPCHK(x)                              \
    do {                                       \
        hipError_t _e = (x);                   \
        if (_e != hipSuccess) return hip_fail(_e); \
    } while (0)

template <typename T>
int upload(void** dst, const std::vector<double>& v) {
    std::vector<T> h(v.begin(), v.end());
    if (!*dst) HIPCHK(hipMalloc(dst, std::max<size_t>(1, h.size()) * sizeof(T)));
    HIPCHK(hipMemcpy(*dst, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return FDMGPU_OK;
}

// twr[m] = cos(2 pi m/N), twi[m] = -sin(2 pi m/N), sp[j] = sin(pi j/N) + 1/2
// (the DST pre-processing weight, rounded once); m, j < N
std::vector<double> dst_tables(int N) {
    std::vector<double> t(3 * (size_t)N);
    for (int m = 0; m < N; m++) {
        const double a = 2.0 * M_PI * m / N;
        t[m] = std::cos(a);
        t[N + m] = -std::sin(a);
        t[2 * N + m] = std::sin(M_PI * m / N) + 0.5;
    }
    return t;
}

// lambda_k = 4/h^2 sin^2(k pi / (2(n+1))), k = 1..n, computed in double and
// rounded to the plan's element type like the reference's init_lm
// (src/lapl_cube.cpp:145-172, src/lapl_rect.cpp:13-22); mode index k-1.
std::vector<double> dirichlet_lambda(int n, double h, bool as_float) {
    std::vector<double> l(n);
    for (int k = 1; k <= n; k++) {
        const double s = std::sin(k * M_PI * 0.5 / (n + 1));
        const double v = 4. / (h * h) * (s * s);
        l[k - 1] = as_float ? (double)(float)v : v;
    }
    return l;
}

// lambda_k = 4/h^2 sin^2(k pi / n), k = 0..n-1 (periodic axis, init_lm with
// tensor_flag::periodic, src/lapl_cube.cpp:145-172): the eigenvalue of
// half-complex coefficient k (cosine k <= n/2, sine of frequency n-k above).
std::vector<double> periodic_lambda(int n, double h, bool as_float) {
    std::vector<double> l(n);
    for (int k = 0; k < n; k++) {
        const double s = std::sin(k * M_PI / n);
        const double v = 4. / (h * h) * (s * s);
        l[k] = as_float ? (double)(float)v : v;
    }
    return l;
}

bool valid_precision(int p) { return p == FDMGPU_F32 || p == FDMGPU_F64 || p == FDMGPU_F32_REFINE64; }

int alloc_tables(fdmgpu_plan_t* P) {
    const bool f32 = P->precision != FDMGPU_F64;
    const std::vector<double> tx = dst_tables(1 << P->logx);
    int rc = f32 ? upload<float>(&P->tab_x, tx) : upload<double>(&P->tab_x, tx);
    if (rc) return rc;
    if (P->logy == P->logx) {
        P->tab_y = P->tab_x;
    } else {
        const std::vector<double> ty = dst_tables(1 << P->logy);
        rc = f32 ? upload<float>(&P->tab_y, ty) : upload<double>(&P->tab_y, ty);
        if (rc) return rc;
    }
    return FDMGPU_OK;
}

}  // namespace

namespace fdmgpu {

void set_error_text(const char* msg) { std::snprintf(g_errbuf, sizeof g_errbuf, "%s", msg); }

static long partial_blocks(const fdmgpu_plan_t* P) {
    // residual_kernel writes residual_partials() per problem, plane_resid_dst_kernel one per z-plane
    return std::max<long>(residual_partials(P->nx, P->ny, P->nz, P->kind == 3), P->nz);
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t work_bytes_shape(int kind, int precision, int nx, int ny, int nz, int count) {
    if (precision != FDMGPU_F32_REFINE64 || count <= 0) return 0;
    const long nblk = std::max<long>(residual_partials(nx, ny, nz, kind == 3), nz);
    const size_t elems = (size_t)nx * ny * nz;
    return (kind == 3 ? 2 : 1) * align256(sizeof(float) * elems * count) +
           align256(sizeof(double) * 2 * nblk * count) + 2 * align256(sizeof(int) * count) +
           align256(sizeof(double) * count);
}

size_t work_bytes(const fdmgpu_plan_t* P, int count) {
    return work_bytes_shape(P->kind, P->precision, P->nx, P->ny, P->nz, count);
}

Work work_carve(const fdmgpu_plan_t* P, int count, void* mem) {
    Work W;
    if (P->precision != FDMGPU_F32_REFINE64 || count <= 0) return W;
    char* c = (char*)mem;
    const long nblk = partial_blocks(P);
    W.r32 = (float*)c;
    c += align256(sizeof(float) * plan_elems(P) * count);
    if (P->kind == 3) {
        W.u32 = (float*)c;
        c += align256(sizeof(float) * plan_elems(P) * count);
    }
    W.partial = (double*)c;
    c += align256(sizeof(double) * 2 * nblk * count);
    W.done = (int*)c;
    c += align256(sizeof(int) * count);
    W.iters = (int*)c;
    c += align256(sizeof(int) * count);
    W.resid = (double*)c;
    W.cap = count;
    return W;
}

void work_free(Work& W) {
    hipFree(W.r32);   // u32 lives in the same allocation
    hipFree(W.partial);
    hipFree(W.done);
    hipFree(W.iters);
    hipFree(W.resid);
    hipFree(W.fcopy);
    W = Work{};
}

int work_reserve(const fdmgpu_plan_t* P, Work& W, int count) {
    if (W.cap >= count && W.r32) return FDMGPU_OK;
    const double* fcopy = W.fcopy;
    const size_t fcb = W.fcopy_bytes;
    int* bad = W.bad;
    hipFree(W.r32);
    hipFree(W.partial);
    hipFree(W.done);
    hipFree(W.iters);
    hipFree(W.resid);
    W = Work{};
    W.fcopy = const_cast<double*>(fcopy);
    W.fcopy_bytes = fcb;
    W.bad = bad;
    const long elems = plan_elems(P);
    const long nblk = partial_blocks(P);
    const size_t r32b = align256(sizeof(float) * elems * count);
    HIPCHK(hipMalloc(&W.r32, P->kind == 3 ? 2 * r32b : r32b));
    if (P->kind == 3) W.u32 = (float*)((char*)W.r32 + r32b);
    HIPCHK(hipMalloc(&W.partial, sizeof(double) * 2 * nblk * count));
    HIPCHK(hipMalloc(&W.done, sizeof(int) * count));
    HIPCHK(hipMalloc(&W.iters, sizeof(int) * count));
    HIPCHK(hipMalloc(&W.resid, sizeof(double) * count));
    W.cap = count;
    return FDMGPU_OK;
}

bool plane_capable(const fdmgpu_plan_t* P) {
    if (P->kind != 3) return false;
    if (P->logx != P->logy)   // rectangular Dirichlet planes: plane_rect_kernel
        return P->periodic == 0 && plane_rect_fits(P->logx, P->logy, P->precision == FDMGPU_F64);
    return P->precision == FDMGPU_F64 ? plane_fits<double>(P->logx) : plane_fits<float>(P->logx);
}

// Square planes (the plane kernels with the fused residual and the streaming variant).
static bool square_planes(const fdmgpu_plan_t* P) { return P->logx == P->logy; }

// Plane kernel of the default variant: the streaming kernel for plain FP64
// planes of N = 128 (one plane fills a CU's LDS), else the plane kernel where
// the plane fits LDS, else the axis tiles.
int default_plane(const fdmgpu_plan_t* P) {
    if (!plane_capable(P)) return 0;
    return square_planes(P) && P->precision == FDMGPU_F64 && plane_stream_ok(P->logx, true) ? 2 : 1;
}

std::vector<Variant> plan_variants(const fdmgpu_plan_t* P) {
    const bool f64 = P->precision == FDMGPU_F64;
    auto tiles = [&](int logn) {
        std::vector<int> t;
        for (int c : {16, 32, 64})
            if (f64 ? tile_ok<double>(logn, c) : tile_ok<float>(logn, c)) t.push_back(c);
        return t;
    };
    std::vector<int> gms{1};
    if (P->kind == 3 && zcr_gmul_ok(P->logz, 2)) gms.push_back(2);
    if (P->kind == 3 && !(P->periodic & FDMGPU_PERIODIC_Z))
        for (int gm : {3, 4, 5})
            if (zcr_gmul_ok(P->logz, gm)) gms.push_back(gm);
    std::vector<Variant> out;
    Variant d;
    d.plane = default_plane(P);
    out.push_back(d);   // the default first
    if (d.plane == 1 && square_planes(P) && !f64 && plane_stream_ok(P->logx, false)) {
        Variant v;   // plane_stream32_kernel with the default z solve
        v.plane = 2;
        out.push_back(v);
    }
    const bool refine = P->precision == FDMGPU_F32_REFINE64;
    for (int gm : gms) {
        if (plane_capable(P) && square_planes(P) && f64 && plane_stream_ok(P->logx, true)) {
            Variant v;   // plane_stream_kernel (plain FP64 planes of N = 128)
            v.plane = 2;
            v.zcr_gmul = gm;
            out.push_back(v);
        }
        if (plane_capable(P) && square_planes(P) && !f64 && plane_stream_ok(P->logx, false)) {
            // plane_stream32_kernel (FP32 planes of N = 128: LaplCube<float>, the
            // refinement's FP32-to-FP32 transforms), with the separate residual pass
            Variant v;
            v.plane = 2;
            v.zcr_gmul = gm;
            out.push_back(v);
        }
        if (plane_capable(P)) {
            // fuse selects the refinement's fused residual pass: plain plans
            // list 0 only, the value of their default
            for (int fu : (refine && square_planes(P) ? std::vector<int>{1, 0} : std::vector<int>{0})) {
                Variant v;
                v.plane = 1;
                v.zcr_gmul = gm;
                v.fuse = fu;
                out.push_back(v);
            }
        }
        for (int tx : (P->kind == 3 || P->fft2 ? tiles(P->logx) : std::vector<int>{0}))
            for (int ty : tiles(P->logy)) {
                if (P->kind == 4 && !(f64 ? tile_ok<double>(P->logx, ty) : tile_ok<float>(P->logx, ty)))
                    continue;   // the phi pass shares tile_y
                Variant v;
                v.plane = 0;
                v.tile_x = tx;
                v.tile_y = ty;
                v.zcr_gmul = gm;
                out.push_back(v);
            }
    }
    return out;
}

bool variant_valid(const fdmgpu_plan_t* P, const Variant& v) {
    for (const Variant& c : plan_variants(P))
        if (c.plane == v.plane && c.tile_x == v.tile_x && c.tile_y == v.tile_y && c.zcr_gmul == v.zcr_gmul &&
            (c.fuse == v.fuse || !c.plane))
            return true;
    return false;
}

}  // namespace fdmgpu

namespace {

// Records HIP events around one launch when timing is enabled.
struct Timed {
    fdmgpu_plan_t* P;
    size_t slot = 0;
    hipStream_t s;
    bool on;
    Timed(fdmgpu_plan_t* P_, int k, hipStream_t s_) : P(P_), s(s_), on((P_->timing >> k) & 1u) {
        if (!on) return;
        slot = P->nrec++;
        while (P->evs.size() < 2 * P->nrec) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) { on = false; P->nrec--; return; }
            P->evs.push_back(e);
        }
        if (P->rec_kernel.size() < P->nrec) P->rec_kernel.resize(P->nrec);
        P->rec_kernel[slot] = k;
        (void)hipEventRecord(P->evs[2 * slot], s);
    }
    ~Timed() {
        if (on) (void)hipEventRecord(P->evs[2 * slot + 1], s);
    }
};

int harvest_timing(fdmgpu_plan_t* P) {
    if (P->nrec == 0) return FDMGPU_OK;
    HIPCHK(hipEventSynchronize(P->evs[2 * (P->nrec - 1) + 1]));
    for (size_t i = 0; i < P->nrec; i++) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, P->evs[2 * i], P->evs[2 * i + 1]));
        P->tot_ms[P->rec_kernel[i]] += ms;
        P->launches[P->rec_kernel[i]] += 1;
    }
    P->nrec = 0;
    return FDMGPU_OK;
}

// LaplCyl3FFT2's phi transforms use 1/sqrt(pi) held in T (SQRT_M_1_PI,
// src/lapl_cyl.h:175): forward factor dphi * SQRT_M_1_PI, inverse SQRT_M_1_PI.
constexpr double kSqrt1Pi = 0.56418958354775629;
template <typename T>
double cyl_phi_scale(const fdmgpu_plan_t* P) {
    return (2 * M_PI / (double)P->nz) * (double)(T)kSqrt1Pi;
}

// ---- the three stages, templated on compute type T ------------------------
// Every stage addresses problem p of a batch at ptr + p * elems, or at tab[p]
// when a device table of per-problem pointers is given (tab != null).
using Tab = const void* const*;
// Forward xy transforms: in (Tin) -> out (T).
template <typename Tin, typename T>
int stage_fwd(fdmgpu_plan_t* P, const void* in, Tab in_tab, void* out, Tab out_tab, int count, hipStream_t s,
              const int* done, int plane_class = FDMGPU_K_XY_FWD) {
    const long elems = (long)P->nx * P->ny * P->nz;
    const double sx = P->dx * std::sqrt(2. / P->lx), sy = P->dy * std::sqrt(2. / P->ly);
    if (P->kind == 4) {   // LaplCyl3FFT2: phi (slowest axis) then z, as "y" passes
        AxisArgs ap{in, out, 0, P->nx * P->ny, 1, elems, elems, count, cyl_phi_scale<T>(P), P->tab_x, s, false, done,
                    kRdftFwd, in_tab, out_tab, P->var.tile_y};
        {
            Timed tm(P, FDMGPU_K_Y_TILE, s);
            HIPCHK((launch_yaxis<Tin, T, T>(P->logx, ap)));
        }
        AxisArgs az{out, out, 0, P->nx, P->nz, elems, elems, count, sy, P->tab_y, s, false, done, P->kfwd, out_tab,
                    out_tab, P->var.tile_y};
        Timed tm(P, FDMGPU_K_Y_TILE, s);
        HIPCHK((launch_yaxis<T, T, T>(P->logy, az)));
        return FDMGPU_OK;
    }
    if (P->kind == 3 && P->var.plane && P->logx == P->logy && plane_fits<T>(P->logx)) {
        PlaneArgs a{in, out, elems, elems, P->nz, count, sx, sy, P->tab_x, s, false, done, P->kfwd, in_tab, out_tab};
        a.stream_planes = P->var.plane == 2;
        Timed tm(P, plane_class, s);
        HIPCHK((launch_plane<Tin, T, T>(P->logx, a)));
        return FDMGPU_OK;
    }
    if (P->kind == 3 && P->var.plane && P->logx != P->logy && P->kfwd == kDst &&
        plane_rect_fits(P->logx, P->logy, sizeof(T) == 8)) {
        PlaneArgs a{in, out, elems, elems, P->nz, count, sx, sy, P->tab_x, s, false, done, P->kfwd, in_tab, out_tab};
        a.tables_y = P->tab_y;
        Timed tm(P, plane_class, s);
        HIPCHK((launch_plane_rect<Tin, T, T>(P->logx, P->logy, a)));
        return FDMGPU_OK;
    }
    if (P->kind == 3) {
        AxisArgs ax{in, out, (long)P->ny * P->nz, P->nx, P->nz, elems, elems, count, sx, P->tab_x, s, false, done,
                    P->kfwd, in_tab, out_tab, P->var.tile_x};
        {
            Timed tm(P, FDMGPU_K_X_TILE, s);
            HIPCHK((launch_xaxis<Tin, T, T>(P->logx, ax)));
        }
        AxisArgs ay{out, out, 0, P->nx, P->nz, elems, elems, count, sy, P->tab_y, s, false, done, P->kfwd, out_tab,
                    out_tab, P->var.tile_y};
        Timed tm(P, FDMGPU_K_Y_TILE, s);
        HIPCHK((launch_yaxis<T, T, T>(P->logy, ay)));
        return FDMGPU_OK;
    }
    // rect: y (LaplRectFFT2: then x)
    AxisArgs ay{in, out, 0, P->nx, 1, elems, elems, count, sy, P->tab_y, s, false, done, P->kfwd, in_tab, out_tab,
                P->var.tile_y};
    {
        Timed tm(P, FDMGPU_K_Y_TILE, s);
        HIPCHK((launch_yaxis<Tin, T, T>(P->logy, ay)));
    }
    if (P->fft2) {
        AxisArgs ax{out, out, (long)P->ny, P->nx, 1, elems, elems, count, sx, P->tab_x, s, false, done, P->kxf,
                    out_tab, out_tab, P->var.tile_x};
        Timed tm(P, FDMGPU_K_X_TILE, s);
        HIPCHK((launch_xaxis<T, T, T>(P->logx, ax)));
    }
    return FDMGPU_OK;
}

// Inverse xy transforms: in (T) -> out (Tout), optionally accumulating; with
// an FP32 `base` (plane path only) out = base + result.
template <typename T, typename Tout>
int stage_inv(fdmgpu_plan_t* P, void* in, Tab in_tab, void* out, Tab out_tab, int count, hipStream_t s, bool accum,
              const int* done, const float* base = nullptr) {
    const long elems = (long)P->nx * P->ny * P->nz;
    const double sx = std::sqrt(2. / P->lx), sy = std::sqrt(2. / P->ly);
    if (P->kind == 4) {   // LaplCyl3FFT2: inverse z, then inverse phi
        if (accum || base) return FDMGPU_ERR_INVALID_ARG;
        AxisArgs az{in, in, 0, P->nx, P->nz, elems, elems, count, sy, P->tab_y, s, false, done, P->kinv, in_tab,
                    in_tab, P->var.tile_y};
        {
            Timed tm(P, FDMGPU_K_Y_TILE, s);
            HIPCHK((launch_yaxis<T, T, T>(P->logy, az)));
        }
        AxisArgs ap{in, out, 0, P->nx * P->ny, 1, elems, elems, count, (double)(T)kSqrt1Pi, P->tab_x, s, false, done,
                    kRdftInv, in_tab, out_tab, P->var.tile_y};
        Timed tm(P, FDMGPU_K_Y_TILE, s);
        HIPCHK((launch_yaxis<T, Tout, T>(P->logx, ap)));
        return FDMGPU_OK;
    }
    if (P->kind == 3 && P->var.plane && P->logx == P->logy && plane_fits<T>(P->logx)) {
        PlaneArgs a{in, out, elems, elems, P->nz, count, sx, sy, P->tab_x, s, accum, done, P->kinv, in_tab, out_tab,
                    base};
        a.stream_planes = P->var.plane == 2;
        Timed tm(P, accum ? FDMGPU_K_XY_INV_ACC : FDMGPU_K_XY_INV, s);
        HIPCHK((launch_plane<T, Tout, T>(P->logx, a)));
        return FDMGPU_OK;
    }
    if (P->kind == 3 && P->var.plane && P->logx != P->logy && P->kinv == kDst &&
        plane_rect_fits(P->logx, P->logy, sizeof(T) == 8)) {
        PlaneArgs a{in, out, elems, elems, P->nz, count, sx, sy, P->tab_x, s, accum, done, P->kinv, in_tab, out_tab,
                    base};
        a.tables_y = P->tab_y;
        Timed tm(P, accum ? FDMGPU_K_XY_INV_ACC : FDMGPU_K_XY_INV, s);
        HIPCHK((launch_plane_rect<T, Tout, T>(P->logx, P->logy, a)));
        return FDMGPU_OK;
    }
    if (base) return FDMGPU_ERR_INVALID_ARG;   // only the plane path adds onto a separate base
    if (P->kind == 3) {
        AxisArgs ay{in, in, 0, P->nx, P->nz, elems, elems, count, sy, P->tab_y, s, false, done, P->kinv, in_tab,
                    in_tab, P->var.tile_y};
        {
            Timed tm(P, FDMGPU_K_Y_TILE, s);
            HIPCHK((launch_yaxis<T, T, T>(P->logy, ay)));
        }
        AxisArgs ax{in, out, (long)P->ny * P->nz, P->nx, P->nz, elems, elems, count, sx, P->tab_x, s, accum, done,
                    P->kinv, in_tab, out_tab, P->var.tile_x};
        Timed tm(P, FDMGPU_K_X_TILE, s);
        HIPCHK((launch_xaxis<T, Tout, T>(P->logx, ax)));
        return FDMGPU_OK;
    }
    if (P->fft2) {   // LaplRectFFT2: inverse x in place first, then y as LaplRect
        AxisArgs ax{in, in, (long)P->ny, P->nx, 1, elems, elems, count, sx, P->tab_x, s, false, done, P->kxi, in_tab,
                    in_tab, P->var.tile_x};
        Timed tm(P, FDMGPU_K_X_TILE, s);
        HIPCHK((launch_xaxis<T, T, T>(P->logx, ax)));
    }
    if (accum) {   // rect: inverse y into the FP32 buffer, then accumulate through an x pass-free add
        AxisArgs ay{in, in, 0, P->nx, 1, elems, elems, count, sy, P->tab_y, s, false, done, P->kinv, in_tab, in_tab,
                    P->var.tile_y};
        Timed tm(P, FDMGPU_K_Y_TILE, s);
        HIPCHK((launch_yaxis<T, T, T>(P->logy, ay)));
        return -1;   // caller adds
    }
    AxisArgs ay{in, out, 0, P->nx, 1, elems, elems, count, sy, P->tab_y, s, false, done, P->kinv, in_tab, out_tab,
                P->var.tile_y};
    Timed tm(P, FDMGPU_K_Y_TILE, s);
    HIPCHK((launch_yaxis<T, Tout, T>(P->logy, ay)));
    return FDMGPU_OK;
}

// FP32 plans (LaplCube<float>, LaplRect<float>) eliminate in double: the
// systems of the lowest modes are nearly singular and an FP32 elimination
// would lose ~cond * eps there, where the reference's z transforms keep
// float accuracy.  The inner solves of FP32+refine stay FP32 (refinement
// recovers the accuracy).
template <typename T>
int stage_tridiag(fdmgpu_plan_t* P, void* data, Tab data_tab, int count, hipStream_t s, const int* done) {
    const long elems = (long)P->nx * P->ny * P->nz;
    if (P->kind == 3) {
        const double b = 1.0 / (P->hz_op * P->hz_op);
        const double gamma = P->dz * (double)(1 << P->logz) / P->lz;
        ZcrArgs z{data, (long)P->nx * P->ny, P->nx, P->lmx, P->lmy, b, -2.0 * b,
                  std::fabs(gamma - 1.0) < 4e-16 ? 1.0 : gamma, elems, count, s, done, P->periodic != 0,
                  P->precision == FDMGPU_F32, data_tab, P->var.zcr_gmul};
        Timed tm(P, FDMGPU_K_ZCR, s);
        HIPCHK(launch_zcr<T>(P->logz, z));
        return FDMGPU_OK;
    }
    if (P->kind == 4) {   // LaplCyl3FFT2: per (phi, z) mode the r system (lm_phi per row group, lm_z per row)
        RowCrArgs r{data, P->nx, P->ny * P->nz, P->lmx, P->d_sc, P->d_Ls, P->d_Us, 1.0 / (P->dx * P->dx), elems,
                    count, s, done, P->precision == FDMGPU_F32, data_tab, P->lmy, P->ny};
        Timed tm(P, FDMGPU_K_ROWCR, s);
        HIPCHK(launch_rowcr<T>(r));
        return FDMGPU_OK;
    }
    if (P->fft2) {
        const bool zm = (P->periodic & FDMGPU_PERIODIC_X) && (P->periodic & FDMGPU_PERIODIC_Y);
        Rect2DivArgs d{data, P->nx, P->ny, P->lmy, P->lmx, P->d_sc, zm, elems, count, s, done, data_tab};
        Timed tm(P, FDMGPU_K_ROWCR, s);
        HIPCHK(launch_rect2_divide<T>(d));
        return FDMGPU_OK;
    }
    RowCrArgs r{data, P->nx, P->ny, P->lmy, P->d_sc, P->d_Ls, P->d_Us, 1.0 / (P->dx * P->dx), elems, count, s, done,
                P->precision == FDMGPU_F32, data_tab};
    r.pivot = P->pivot_rows;
    Timed tm(P, FDMGPU_K_ROWCR, s);
    HIPCHK(launch_rowcr<T>(r));
    return FDMGPU_OK;
}

template <typename T>
int solve_plain(fdmgpu_plan_t* P, void* ans, Tab ans_tab, const void* rhs, Tab rhs_tab, int count, hipStream_t s) {
    int rc = stage_fwd<T, T>(P, rhs, rhs_tab, ans, ans_tab, count, s, nullptr);
    if (rc) return rc;
    rc = stage_tridiag<T>(P, ans, ans_tab, count, s, nullptr);
    if (rc) return rc;
    return stage_inv<T, T>(P, ans, ans_tab, ans, ans_tab, count, s, false, nullptr);
}

__global__ void add_f32_to_f64(const BatchPtr<double> u, const float* __restrict__ c, const long n, const long bs,
                               const int* __restrict__ done) {
    if (done && done[blockIdx.y]) return;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) as_global(u(blockIdx.y))[i] += (double)c[blockIdx.y * bs + i];
}

// FP32 inner solves + FP64 residual/update/norm.
//   solve 1:      c = solve32(f), u = c
//   solve i >= 2: r = f - A u (FP64, fused into the forward plane DST for
//                 cubes), ||r||/||f|| -> per-problem flag, c = solve32(r), u += c
//   finally:      ||f - A u||/||f|| of the last iterate.
// On the plane path the iterate after solve 1 holds FP32 values, so when a
// second solve follows in the same call it stays in the FP32 scratch u32
// (4 bytes per point written and read twice instead of 8); solve 2 writes
// u = u32 + c into ans, and problems that converged on solve 1 get u = u32.
// A problem whose residual reached rtol is flagged on the device; every later
// launch skips it, so the host never synchronises inside a solve.
// first > 1 continues a refinement whose iterate (in ans) after `first - 1`
// inner solves did not reach the tolerance; last = the final inner solve.
// Periodic cubes (every axis periodic; singular operator): the zero-mean
// solution of A u = f - mean f, as the reference's periodic solve returns it
// (its (0,0,0) mode is zeroed, src/lapl_cube.cpp:82-84).
//   once per call: mean f per problem (fsum + fmean)
//   solve 1:      r32 = f - mean f, c = solve32(r32), u = c - mean c
//   solve i >= 2: c = solve32(r32), u += c - mean c
//   after each:   r = (f - mean f) - A u (wrap-around stencil), ||r|| /
//                 ||f - mean f|| -> per-problem flag; r32 = r for the next
// The inner solve is the plan's FP32 periodic solve in place on r32; the
// per-problem means of f and of the correction live in the FP32 iterate
// scratch u32, which the periodic path does not otherwise use.
int solve_refine_periodic(fdmgpu_plan_t* P, Work& W, double* ans, Tab ans_tab, const double* rhs, Tab rhs_tab,
                          int count, hipStream_t s, int first, int last) {
    const long elems = (long)P->nx * P->ny * P->nz;
    PeriodicResidualArgs pa{};
    pa.f = rhs;
    pa.f_tab = rhs_tab;
    pa.u = ans;
    pa.u_tab = ans_tab;
    pa.fmean = (double*)W.u32;
    pa.cmean = (double*)W.u32 + count;   // u32 holds 4 elems >= 16 bytes per problem (n >= 2 per axis)
    pa.r32 = W.r32;
    pa.partial = W.partial;
    pa.nx = P->nx;
    pa.ny = P->ny;
    pa.nz = P->nz;
    pa.nblk = (int)std::max<long>(1, std::min<long>(residual_partials(P->nx, P->ny, P->nz, true), elems));
    pa.ix = 1.0 / (P->hx_op * P->hx_op);
    pa.iy = 1.0 / (P->dy * P->dy);
    pa.iz = 1.0 / (P->hz_op * P->hz_op);
    pa.bs = elems;
    pa.batch = count;
    pa.tol = P->rtol;
    pa.done = W.done;
    pa.iters = W.iters;
    pa.resid = W.resid;
    pa.stream = s;
    {
        Timed tm(P, FDMGPU_K_OTHER, s);
        HIPCHK(launch_periodic_fmean(pa));
    }
    if (first > 1) {   // the current iterate: flags, and r32 for the next solve
        pa.check = true;
        pa.it = first - 1;
        Timed tm(P, FDMGPU_K_RESIDUAL_RHS, s);
        HIPCHK(launch_periodic_residual(pa));
    }
    for (int it = first; it <= last; ++it) {
        if (it == 1) {   // r32 = f - mean f
            pa.u = nullptr;
            pa.u_tab = nullptr;
            pa.r32 = W.r32;
            pa.check = false;
            Timed tm(P, FDMGPU_K_RESIDUAL_RHS, s);
            HIPCHK(launch_periodic_residual(pa));
        }
        int rc = stage_fwd<float, float>(P, W.r32, nullptr, W.r32, nullptr, count, s, W.done, FDMGPU_K_XY_FWD_R32);
        if (rc) return rc;
        rc = stage_tridiag<float>(P, W.r32, nullptr, count, s, W.done);
        if (rc) return rc;
        rc = stage_inv<float, float>(P, W.r32, nullptr, W.r32, nullptr, count, s, false, W.done);
        if (rc) return rc;
        {
            Timed tm(P, FDMGPU_K_OTHER, s);
            HIPCHK(launch_periodic_update(pa, ans, ans_tab, W.r32, it == 1));
        }
        pa.u = ans;
        pa.u_tab = ans_tab;
        pa.r32 = it < last ? W.r32 : nullptr;
        pa.check = true;
        pa.it = it;
        pa.bad = it == last ? W.bad : nullptr;
        Timed tm(P, pa.r32 ? FDMGPU_K_RESIDUAL_RHS : FDMGPU_K_RESIDUAL, s);
        HIPCHK(launch_periodic_residual(pa));
    }
    return FDMGPU_OK;
}

int solve_refine(fdmgpu_plan_t* P, Work& W, double* ans, Tab ans_tab, const double* rhs, Tab rhs_tab,
                 int count, hipStream_t s, int first, int last) {
    int rc = work_reserve(P, W, count);
    if (rc) return rc;
    const long elems = (long)P->nx * P->ny * P->nz;
    if (!ans_tab && (const void*)ans == (const void*)rhs) {   // f is needed by every residual: keep a copy
        const size_t bytes = sizeof(double) * elems * count;
        if (W.fcopy_bytes < bytes) {
            hipFree(W.fcopy);
            W.fcopy = nullptr;
            W.fcopy_bytes = 0;
            HIPCHK(hipMalloc(&W.fcopy, bytes));
            W.fcopy_bytes = bytes;
        }
        Timed tm(P, FDMGPU_K_OTHER, s);
        HIPCHK(hipMemcpyAsync(W.fcopy, rhs, bytes, hipMemcpyDeviceToDevice, s));
        rhs = W.fcopy;
    }
    {
        Timed tm(P, FDMGPU_K_OTHER, s);
        HIPCHK(launch_fill_int(W.done, W.iters, count, 0, s));
    }
    if (P->periodic) return solve_refine_periodic(P, W, ans, ans_tab, rhs, rhs_tab, count, s, first, last);
    ResidualArgs ra{};
    ra.f = rhs;
    ra.u = ans;
    ra.r32 = W.r32;
    ra.partial = W.partial;
    ra.nx = P->nx;
    ra.ny = P->ny;
    ra.nz = P->nz;
    ra.three = P->kind == 3;
    ra.ix = 1.0 / (P->hx_op * P->hx_op);
    ra.iy = 1.0 / (P->dy * P->dy);
    ra.iz = P->kind == 3 ? 1.0 / (P->hz_op * P->hz_op) : 0.0;
    ra.sy = P->scaled ? P->d_sc64 : nullptr;
    ra.lx = P->d_Ls64;
    ra.ux = P->d_Us64;
    ra.bs = elems;
    ra.batch = count;
    ra.tol = P->rtol;
    ra.done = W.done;
    ra.iters = W.iters;
    ra.resid = W.resid;
    ra.stream = s;
    ra.f_tab = rhs_tab;
    ra.u_tab = ans_tab;
    const bool fused = plane_capable(P) && square_planes(P) && P->var.plane && P->var.fuse;
    // u1 in FP32 scratch: separate-residual plane path (Dirichlet planes), solves 1 and 2 in
    // this call (the fused forward pass reads an FP32 iterate slower than it saves)
    const bool u1f = plane_capable(P) && P->var.plane && !fused && P->kinv == kDst && W.u32 && first == 1 &&
                     last >= 2;

    // Fused path (cubes whose planes fit LDS): the forward DST of solve i >= 2
    // computes r = f - A u itself; convergence of the iterate after solve i is
    // certified by a norm-only FP64 residual pass.  Other shapes: the residual
    // pass writes r32 (FP32) for the next solve and certifies at the same time.
    if (first > 1 && !fused) {   // the next solve's rhs: r32 = f - A u of the current iterate
        ra.it = first - 1;
        ra.r32 = W.r32;
        Timed tm(P, FDMGPU_K_RESIDUAL_RHS, s);
        HIPCHK(launch_residual(ra));
    }
    for (int it = first; it <= last; ++it) {
        if (it == 1) {
            rc = stage_fwd<double, float>(P, rhs, rhs_tab, W.r32, nullptr, count, s, W.done);
        } else if (fused) {
            PlaneResidArgs pa{rhs, ans, W.r32, W.partial, elems, P->nz, count, ra.ix, ra.iy, ra.iz,
                              P->dx * std::sqrt(2. / P->lx), P->dy * std::sqrt(2. / P->ly), P->tab_x, s, W.done,
                              rhs_tab, ans_tab};
            {
                Timed tm(P, FDMGPU_K_XY_FWD_RESID, s);
                HIPCHK(launch_plane_resid<double>(P->logx, pa));
            }
            if (it == 2 && first == 1) {   // the iterate after solve 1 was not certified
                Timed tm(P, FDMGPU_K_RESIDUAL, s);
                HIPCHK(launch_check(W.partial, P->nz, P->rtol, W.done, W.iters, W.resid, 1, count, s));
            }
            rc = 0;
        } else {
            rc = stage_fwd<float, float>(P, W.r32, nullptr, W.r32, nullptr, count, s, W.done, FDMGPU_K_XY_FWD_R32);
        }
        if (rc) return rc;
        rc = stage_tridiag<float>(P, W.r32, nullptr, count, s, W.done);
        if (rc) return rc;
        if (u1f && it == 1)
            rc = stage_inv<float, float>(P, W.r32, nullptr, W.u32, nullptr, count, s, false, W.done);
        else
            rc = stage_inv<float, double>(P, W.r32, nullptr, ans, ans_tab, count, s, it > 1, W.done,
                                          u1f && it == 2 ? W.u32 : nullptr);
        if (rc == -1) {   // rect accumulate
            dim3 g((unsigned)((elems + 255) / 256), count);
            Timed tm(P, FDMGPU_K_OTHER, s);
            hipLaunchKernelGGL(add_f32_to_f64, g, dim3(256), 0, s, BatchPtr<double>{ans, ans_tab, elems},
                               (const float*)W.r32, elems, elems, (const int*)W.done);
            HIPCHK(hipGetLastError());
            rc = 0;
        }
        if (rc) return rc;
        if (fused && it == 1 && last > 1) continue;   // solve 2's fused pass checks it
        ra.it = it;
        ra.r32 = (fused || it == last) ? nullptr : W.r32;
        ra.u = u1f && it == 1 ? (const double*)W.u32 : ans;
        ra.u_tab = u1f && it == 1 ? nullptr : ans_tab;
        ra.u_f32 = u1f && it == 1;
        ra.bad = it == last ? W.bad : nullptr;
        Timed tm(P, ra.r32 ? FDMGPU_K_RESIDUAL_RHS : FDMGPU_K_RESIDUAL, s);
        HIPCHK(launch_residual(ra));
    }
    return FDMGPU_OK;
}

int solve_device_impl(fdmgpu_plan_t* P, void* ans, Tab ans_tab, const void* rhs, Tab rhs_tab, int count,
                      hipStream_t s) {
    P->last_stream = s;
    P->last_count = count;
    const int rc = plan_solve(P, P->w, ans, ans_tab, rhs, rhs_tab, count, s);
    P->last_iters = P->w.iters;
    P->last_resid = P->w.resid;
    return rc;
}

}  // namespace

int fdmgpu::plan_solve(fdmgpu_plan_t* P, Work& W, void* ans, Tab ans_tab, const void* rhs, Tab rhs_tab, int count,
                       hipStream_t s) {
    return plan_solve_iters(P, W, ans, ans_tab, rhs, rhs_tab, count, s, 1, P->max_iter);
}

int fdmgpu::plan_solve_iters(fdmgpu_plan_t* P, Work& W, void* ans, Tab ans_tab, const void* rhs, Tab rhs_tab,
                             int count, hipStream_t s, int first, int last) {
    HIPCHK(hipSetDevice(P->device));
    if (count <= 0) return FDMGPU_OK;
    // the launchers check hipGetLastError(); drop a stale error of an earlier,
    // unrelated runtime call on this thread so it is not reported as ours
    (void)hipGetLastError();
    switch (P->precision) {
        case FDMGPU_F32: return solve_plain<float>(P, ans, ans_tab, rhs, rhs_tab, count, s);
        case FDMGPU_F64: return solve_plain<double>(P, ans, ans_tab, rhs, rhs_tab, count, s);
        default:
            if (first < 1 || last < first || last > 64) return FDMGPU_ERR_INVALID_ARG;
            return solve_refine(P, W, (double*)ans, ans_tab, (const double*)rhs, rhs_tab, count, s, first, last);
    }
}

namespace {

int common_init(fdmgpu_plan_t* P, int precision) {
    P->precision = precision;
    P->elem = precision == FDMGPU_F32 ? 4 : 8;
    P->csize = precision == FDMGPU_F64 ? 8 : 4;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) {
        std::snprintf(g_errbuf, sizeof g_errbuf, "no HIP device available");
        return FDMGPU_ERR_DEVICE;
    }
    HIPCHK(hipGetDevice(&P->device));
    return FDMGPU_OK;
}

void free_plan(fdmgpu_plan_t* P) {
    if (!P) return;
    if (P->tab_y != P->tab_x) hipFree(P->tab_y);
    hipFree(P->tab_x);
    hipFree(P->lmx);
    hipFree(P->lmy);
    hipFree(P->d_sc);
    hipFree(P->d_Ls);
    hipFree(P->d_Us);
    hipFree(P->d_sc64);
    hipFree(P->d_Ls64);
    hipFree(P->d_Us64);
    work_free(P->w);
    hipFree(P->stage);
    for (hipEvent_t e : P->evs) (void)hipEventDestroy(e);
    if (P->own) hipStreamDestroy(P->own);
    delete P;
}

int set_scales_fft2(fdmgpu_plan_t* P, const double* s) {
    const int nx = P->nx;
    const bool xper = (P->periodic & FDMGPU_PERIODIC_X) != 0, yper = (P->periodic & FDMGPU_PERIODIC_Y) != 0;
    const int x1 = xper ? 0 : 1;
    if (s && P->precision == FDMGPU_F32_REFINE64) return FDMGPU_ERR_UNSUPPORTED;   // no real-space stencil
    P->h_sc.assign(nx, 1.0);
    for (int j = 0; j < nx; j++) {
        if (s) P->h_sc[j] = s[j + x1];
        if (P->precision == FDMGPU_F32) P->h_sc[j] = (double)(float)P->h_sc[j];
    }
    // every mode's divisor must be nonzero (the double-periodic zero mode is replaced)
    const std::vector<double> ly = yper ? periodic_lambda(P->ny, P->dy, false) : dirichlet_lambda(P->ny, P->dy, false);
    const std::vector<double> lx = P->h_lmx;
    for (int k = 0; k < P->ny; k++)
        for (int j = 0; j < nx; j++)
            if (!(k == 0 && j == 0 && xper && yper) && ly[k] * P->h_sc[j] + lx[j] == 0.0) {
                std::snprintf(g_errbuf, sizeof g_errbuf, "LaplRectFFT2 mode (%d, %d) has a zero eigenvalue", k, j);
                return FDMGPU_ERR_SINGULAR;
            }
    const bool f32 = P->precision != FDMGPU_F64;
    int rc;
    if ((rc = f32 ? upload<float>(&P->d_sc, P->h_sc) : upload<double>(&P->d_sc, P->h_sc))) return rc;
    return upload<double>((void**)&P->d_sc64, P->h_sc);
}

// The pivot sequence of ?gtsv on one tridiagonal system (coefficients only,
// overwritten): the 1-based column of the first zero pivot, 0 if none.
int gtsv_zero_pivot(const int n, double* dl, double* d, double* du) {
    for (int i = 0; i < n - 1; i++) {
        if (std::fabs(d[i]) >= std::fabs(dl[i])) {
            if (d[i] == 0.0) return i + 1;
            const double fact = dl[i] / d[i];
            d[i + 1] -= fact * du[i];
        } else {
            const double fact = d[i] / dl[i];
            d[i] = dl[i];
            const double temp = d[i + 1];
            d[i + 1] = du[i] - fact * temp;
            if (i < n - 2) du[i + 1] = -fact * du[i + 1];
            du[i] = temp;
        }
    }
    return n > 0 && d[n - 1] == 0.0 ? n : 0;
}

int set_scales_impl(fdmgpu_plan_t* P, const double* s, const double* L, const double* U) {
    if (P->fft2) return set_scales_fft2(P, s);
    const int nx = P->nx;
    P->h_sc.assign(nx, 1.0);
    P->h_Ls.assign(nx, 1.0);
    P->h_Us.assign(nx, 1.0);
    for (int j = 0; j < nx; j++) {
        if (s) P->h_sc[j] = s[j + 1];
        if (L) P->h_Ls[j] = L[j + 1];
        if (U) P->h_Us[j] = U[j + 1];
    }
    if (P->precision == FDMGPU_F32) {   // LaplRect<float> stores the scales as float
        for (int j = 0; j < nx; j++) {
            P->h_sc[j] = (double)(float)P->h_sc[j];
            P->h_Ls[j] = (double)(float)P->h_Ls[j];
            P->h_Us[j] = (double)(float)P->h_Us[j];
        }
    }
    // Row systems that are all diagonally dominant (the default scales, the
    // cylindrical ones) are solved by cyclic reduction; otherwise every row
    // goes through Gaussian elimination with partial pivoting, the reference's
    // ?gtsv.  A row whose elimination meets a zero pivot is singular: the
    // reference aborts there (verify(info == 0), src/lapl_rect.cpp:91).
    std::vector<double> lam = P->periodic ? periodic_lambda(P->ny, P->dy, false) : dirichlet_lambda(P->ny, P->dy, false);
    const double i2 = 1.0 / (P->dx * P->dx);
    bool pivot = false, strict = true;   // some row not dominant / every row strictly dominant
    for (int k = 0; k < P->ny && !pivot; k++) {
        for (int j = 0; j < nx; j++) {
            const double d = std::fabs(-2.0 * i2 - lam[k] * P->h_sc[j]);
            const double off = (j > 0 ? std::fabs(P->h_Ls[j]) : 0.0) * i2 + (j < nx - 1 ? std::fabs(P->h_Us[j]) : 0.0) * i2;
            if (!(d > off)) strict = false;
            if (!(d >= off) || d == 0.0) {
                pivot = true;
                break;
            }
        }
    }
    if (!strict) {   // not strictly dominant: a row may be singular
        std::vector<double> dl(nx), d(nx), du(nx);
        for (int k = 0; k < P->ny; k++) {
            for (int j = 0; j < nx; j++) {
                d[j] = -2.0 * i2 - lam[k] * P->h_sc[j];
                dl[j] = j + 1 < nx ? P->h_Ls[j + 1] * i2 : 0.0;
                du[j] = j + 1 < nx ? P->h_Us[j] * i2 : 0.0;
            }
            const int info = gtsv_zero_pivot(nx, dl.data(), d.data(), du.data());
            if (info) {
                std::snprintf(g_errbuf, sizeof g_errbuf,
                              "LaplRect row %d is singular (gtsv info %d: zero pivot in column %d)", k + 1, info, info);
                return FDMGPU_ERR_SINGULAR;
            }
        }
    }
    P->pivot_rows = pivot;
    const bool f32 = P->precision != FDMGPU_F64;
    int rc;
    if ((rc = f32 ? upload<float>(&P->d_sc, P->h_sc) : upload<double>(&P->d_sc, P->h_sc))) return rc;
    if ((rc = f32 ? upload<float>(&P->d_Ls, P->h_Ls) : upload<double>(&P->d_Ls, P->h_Ls))) return rc;
    if ((rc = f32 ? upload<float>(&P->d_Us, P->h_Us) : upload<double>(&P->d_Us, P->h_Us))) return rc;
    if ((rc = upload<double>((void**)&P->d_sc64, P->h_sc))) return rc;
    if ((rc = upload<double>((void**)&P->d_Ls64, P->h_Ls))) return rc;
    if ((rc = upload<double>((void**)&P->d_Us64, P->h_Us))) return rc;
    P->scaled = false;
    for (int j = 0; j < nx; j++)
        if (P->h_sc[j] != 1.0 || P->h_Ls[j] != 1.0 || P->h_Us[j] != 1.0) P->scaled = true;
    return FDMGPU_OK;
}

}  // namespace

extern "C" {

int fdmgpu_version(void) { return 100; }

const char* fdmgpu_strerror(int code) {
    switch (code) {
        case FDMGPU_OK: return "ok";
        case FDMGPU_ERR_INVALID_SIZE: return "transform length must be a power of two (verify((1<<n)==N))";
        case FDMGPU_ERR_INVALID_ARG: return "invalid argument";
        case FDMGPU_ERR_UNSUPPORTED: return "configuration not supported on the GPU";
        case FDMGPU_ERR_OUT_OF_MEMORY: return "device out of memory";
        case FDMGPU_ERR_DEVICE: return g_errbuf[0] ? g_errbuf : "HIP device error";
        case FDMGPU_ERR_SINGULAR: return g_errbuf[0] ? g_errbuf : "tridiagonal system needs pivoting (gtsv info != 0)";
        case FDMGPU_ERR_NOT_CONVERGED: return "refinement did not reach the requested residual";
        case FDMGPU_ERR_ADMISSION: return "request exceeds the HBM budget";
        case FDMGPU_ERR_TOO_LARGE: return g_errbuf[0] ? g_errbuf : "solve needs more than the static scratch";
        default: return "unknown error";
    }
}

int fdmgpu_plan_cube(fdmgpu_plan* plan, int precision, int periodic_mask, double dx, double dy, double dz,
                     double lx, double ly, double lz, int nx, int ny, int nz) {
    if (!plan || !valid_precision(precision)) return FDMGPU_ERR_INVALID_ARG;
    *plan = nullptr;
    if (periodic_mask != 0 && periodic_mask != 7) return FDMGPU_ERR_INVALID_ARG;
    if (nx < 1 || ny < 1 || nz < 1) return FDMGPU_ERR_INVALID_SIZE;
    const bool per = periodic_mask != 0;
    // transform lengths: n+1 on Dirichlet axes, n on periodic ones (src/lapl_cube.h:66-76)
    const int e = per ? 0 : 1;
    const int lgx = ilog2_exact(nx + e), lgy = ilog2_exact(ny + e), lgz = ilog2_exact(nz + e);
    if (lgx < 0 || lgy < 0 || lgz < 0) return FDMGPU_ERR_INVALID_SIZE;
    if (lgx > 9 || lgy > 9 || lgz > 9 || lgx < 1 || lgy < 1 || lgz < 1) return FDMGPU_ERR_UNSUPPORTED;
    fdmgpu_plan_t* P = new (std::nothrow) fdmgpu_plan_t();
    if (!P) return FDMGPU_ERR_OUT_OF_MEMORY;
    int rc = common_init(P, precision);
    if (rc) { free_plan(P); return rc; }
    P->kind = 3;
    P->periodic = periodic_mask;
    P->kfwd = per ? kRdftFwd : kDst;
    P->kinv = per ? kRdftInv : kDst;
    P->nx = nx; P->ny = ny; P->nz = nz;
    P->logx = lgx; P->logy = lgy; P->logz = lgz;
    P->dx = dx; P->dy = dy; P->dz = dz; P->lx = lx; P->ly = ly; P->lz = lz;
    if (precision == FDMGPU_F32_REFINE64) {
        // refinement converges to A^{-1} f; the reference returns g A^{-1} f with
        // g = prod_a h_a N_a / l_a (N = n + 1 Dirichlet, n periodic), which is 1
        // for the usual l = N h geometry
        const double g = dx * (nx + e) / lx * dy * (ny + e) / ly * dz * (nz + e) / lz;
        if (std::fabs(g - 1.0) > 1e-13) { free_plan(P); return FDMGPU_ERR_UNSUPPORTED; }
    }
    // eigenvalue aliasing of the reference (src/lapl_cube.cpp:162,171)
    P->hx_op = (nx == ny) ? dy : dx;
    P->hz_op = (nz == ny) ? dy : dz;
    const bool f32 = precision != FDMGPU_F64;
    const bool tf = precision == FDMGPU_F32;   // the reference's lambda tables are stored as T
    auto lambda = [&](int n, double h) { return per ? periodic_lambda(n, h, tf) : dirichlet_lambda(n, h, tf); };
    std::vector<double> lmy = lambda(ny, dy);
    std::vector<double> lmx = (nx == ny) ? lmy : lambda(nx, dx);
    if ((rc = alloc_tables(P)) ||
        (rc = f32 ? upload<float>(&P->lmx, lmx) : upload<double>(&P->lmx, lmx)) ||
        (rc = f32 ? upload<float>(&P->lmy, lmy) : upload<double>(&P->lmy, lmy))) {
        free_plan(P);
        return rc;
    }
    P->var.plane = default_plane(P);   // the default variant
    apply_tuned(P);
    *plan = P;
    return FDMGPU_OK;
}

int fdmgpu_plan_rect(fdmgpu_plan* plan, int precision, int periodic_mask, double dx, double dy, double lx,
                     double ly, int nx, int ny) {
    if (!plan || !valid_precision(precision)) return FDMGPU_ERR_INVALID_ARG;
    *plan = nullptr;
    if (periodic_mask != 0 && periodic_mask != FDMGPU_PERIODIC_Y) return FDMGPU_ERR_INVALID_ARG;
    if (nx < 1 || ny < 1) return FDMGPU_ERR_INVALID_SIZE;
    const bool per = periodic_mask != 0;
    const int lgy = ilog2_exact(per ? ny : ny + 1);
    if (lgy < 0) return FDMGPU_ERR_INVALID_SIZE;
    if (lgy < 1 || lgy > 9 || nx > 4096) return FDMGPU_ERR_UNSUPPORTED;   // row CR stages 4*nx values in LDS
    if (per && precision == FDMGPU_F32_REFINE64) return FDMGPU_ERR_UNSUPPORTED;
    fdmgpu_plan_t* P = new (std::nothrow) fdmgpu_plan_t();
    if (!P) return FDMGPU_ERR_OUT_OF_MEMORY;
    int rc = common_init(P, precision);
    if (rc) { free_plan(P); return rc; }
    P->kind = 2;
    P->periodic = periodic_mask;
    P->kfwd = per ? kRdftFwd : kDst;
    P->kinv = per ? kRdftInv : kDst;
    P->nx = nx; P->ny = ny; P->nz = 1;
    P->logx = lgy; P->logy = lgy; P->logz = 0;
    P->dx = dx; P->dy = dy; P->lx = lx; P->ly = ly;
    P->hx_op = dx;
    if (precision == FDMGPU_F32_REFINE64 && std::fabs(dy * (ny + 1) / ly - 1.0) > 1e-13) {
        free_plan(P);
        return FDMGPU_ERR_UNSUPPORTED;
    }
    const bool f32 = precision != FDMGPU_F64;
    std::vector<double> lmy = per ? periodic_lambda(ny, dy, precision == FDMGPU_F32)
                                  : dirichlet_lambda(ny, dy, precision == FDMGPU_F32);
    if ((rc = alloc_tables(P)) || (rc = f32 ? upload<float>(&P->lmy, lmy) : upload<double>(&P->lmy, lmy)) ||
        (rc = set_scales_impl(P, nullptr, nullptr, nullptr))) {
        free_plan(P);
        return rc;
    }
    P->var.plane = default_plane(P);   // the default variant
    apply_tuned(P);
    *plan = P;
    return FDMGPU_OK;
}

int fdmgpu_plan_rect_fft2(fdmgpu_plan* plan, int precision, int periodic_mask, double dx, double dy, double lx,
                          double ly, int nx, int ny) {
    if (!plan || !valid_precision(precision)) return FDMGPU_ERR_INVALID_ARG;
    *plan = nullptr;
    if (periodic_mask & ~(FDMGPU_PERIODIC_X | FDMGPU_PERIODIC_Y)) return FDMGPU_ERR_INVALID_ARG;
    if (nx < 1 || ny < 1) return FDMGPU_ERR_INVALID_SIZE;
    const bool yper = periodic_mask & FDMGPU_PERIODIC_Y, xper = periodic_mask & FDMGPU_PERIODIC_X;
    const int lgy = ilog2_exact(yper ? ny : ny + 1), lgx = ilog2_exact(xper ? nx : nx + 1);
    if (lgy < 0 || lgx < 0) return FDMGPU_ERR_INVALID_SIZE;
    if (lgy < 1 || lgy > 9 || lgx < 1 || lgx > 9) return FDMGPU_ERR_UNSUPPORTED;
    if (precision == FDMGPU_F32_REFINE64 &&
        (periodic_mask || std::fabs(dy * (ny + 1) / ly - 1.0) > 1e-13 || std::fabs(dx * (nx + 1) / lx - 1.0) > 1e-13))
        return FDMGPU_ERR_UNSUPPORTED;   // the residual stencil is the transforms' operator only then
    fdmgpu_plan_t* P = new (std::nothrow) fdmgpu_plan_t();
    if (!P) return FDMGPU_ERR_OUT_OF_MEMORY;
    int rc = common_init(P, precision);
    if (rc) { free_plan(P); return rc; }
    P->kind = 2;
    P->fft2 = true;
    P->periodic = periodic_mask;
    P->kfwd = yper ? kRdftFwd : kDst;
    P->kinv = yper ? kRdftInv : kDst;
    P->kxf = xper ? kRdftFwd : kDst;
    P->kxi = xper ? kRdftInv : kDst;
    P->nx = nx; P->ny = ny; P->nz = 1;
    P->logx = lgx; P->logy = lgy; P->logz = 0;
    P->dx = dx; P->dy = dy; P->lx = lx; P->ly = ly;
    const bool as_float = precision == FDMGPU_F32;
    std::vector<double> lmy = yper ? periodic_lambda(ny, dy, as_float) : dirichlet_lambda(ny, dy, as_float);
    std::vector<double> lmx = xper ? periodic_lambda(nx, dx, as_float) : dirichlet_lambda(nx, dx, as_float);
    // lm_x aliases lm_y in the all-Dirichlet instantiation when the point
    // counts match (src/lapl_rect.cpp:36-40)
    const bool alias = !periodic_mask && nx == ny;
    if (alias) lmx = lmy;
    P->hx_op = alias ? dy : dx;
    P->h_lmx = lmx;
    const bool f32 = precision != FDMGPU_F64;
    if ((rc = alloc_tables(P)) || (rc = f32 ? upload<float>(&P->lmy, lmy) : upload<double>(&P->lmy, lmy)) ||
        (rc = f32 ? upload<float>(&P->lmx, lmx) : upload<double>(&P->lmx, lmx)) ||
        (rc = set_scales_impl(P, nullptr, nullptr, nullptr))) {
        free_plan(P);
        return rc;
    }
    P->var.plane = 0;
    apply_tuned(P);
    *plan = P;
    return FDMGPU_OK;
}

int fdmgpu_plan_cyl3_fft2(fdmgpu_plan* plan, int precision, int periodic_mask, double dr, double dz, double r0,
                          double lr, double lz, int nr, int nz, int nphi) {
    if (!plan || !valid_precision(precision)) return FDMGPU_ERR_INVALID_ARG;
    *plan = nullptr;
    if (periodic_mask != 0 && periodic_mask != FDMGPU_PERIODIC_Z) return FDMGPU_ERR_INVALID_ARG;
    if (nr < 1 || nz < 1 || nphi < 1 || !(dr > 0) || !(dz > 0)) return FDMGPU_ERR_INVALID_SIZE;
    const bool zper = periodic_mask != 0;
    const int lgp = ilog2_exact(nphi), lgz = ilog2_exact(zper ? nz : nz + 1);
    if (lgp < 0 || lgz < 0) return FDMGPU_ERR_INVALID_SIZE;
    if (lgp < 1 || lgp > 9 || lgz < 1 || lgz > 9 || nr > 4096) return FDMGPU_ERR_UNSUPPORTED;
    if (precision == FDMGPU_F32_REFINE64) return FDMGPU_ERR_UNSUPPORTED;
    for (int j = 1; j <= nr; j++)
        if (!(r0 + j * dr > 0)) return FDMGPU_ERR_INVALID_ARG;   // r appears as 1/r and 1/r^2
    fdmgpu_plan_t* P = new (std::nothrow) fdmgpu_plan_t();
    if (!P) return FDMGPU_ERR_OUT_OF_MEMORY;
    int rc = common_init(P, precision);
    if (rc) { free_plan(P); return rc; }
    // array [nphi][nz][nr]: x = r (rows of the tridiagonal), y = z, z = phi
    P->kind = 4;
    P->periodic = periodic_mask;
    P->kfwd = zper ? kRdftFwd : kDst;   // z axis
    P->kinv = zper ? kRdftInv : kDst;
    P->nx = nr; P->ny = nz; P->nz = nphi;
    P->logx = lgp; P->logy = lgz; P->logz = 0;   // tab_x: phi tables, tab_y: z tables
    P->dx = dr; P->dy = dz; P->lx = lr; P->ly = lz;
    const bool as_float = precision == FDMGPU_F32;
    const std::vector<double> lm_phi = periodic_lambda(nphi, 2 * M_PI / nphi, as_float);   // 4/dphi^2 sin^2(i dphi/2)
    const std::vector<double> lm_z = zper ? periodic_lambda(nz, dz, as_float) : dirichlet_lambda(nz, dz, as_float);
    std::vector<double> sc(nr), Ls(nr), Us(nr);
    for (int j = 1; j <= nr; j++) {
        const double r = r0 + j * dr;
        sc[j - 1] = 1.0 / (r * r);
        Ls[j - 1] = (r - 0.5 * dr) / r;
        Us[j - 1] = (r + 0.5 * dr) / r;
    }
    const bool f32 = precision != FDMGPU_F64;
    if ((rc = alloc_tables(P)) || (rc = f32 ? upload<float>(&P->lmx, lm_phi) : upload<double>(&P->lmx, lm_phi)) ||
        (rc = f32 ? upload<float>(&P->lmy, lm_z) : upload<double>(&P->lmy, lm_z)) ||
        (rc = f32 ? upload<float>(&P->d_sc, sc) : upload<double>(&P->d_sc, sc)) ||
        (rc = f32 ? upload<float>(&P->d_Ls, Ls) : upload<double>(&P->d_Ls, Ls)) ||
        (rc = f32 ? upload<float>(&P->d_Us, Us) : upload<double>(&P->d_Us, Us))) {
        free_plan(P);
        return rc;
    }
    P->var.plane = 0;
    apply_tuned(P);
    *plan = P;
    return FDMGPU_OK;
}

int fdmgpu_rect_set_scales(fdmgpu_plan plan, const double* lm_y_scale, const double* L_scale,
                           const double* U_scale) {
    if (!plan || plan->kind != 2) return FDMGPU_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(plan->device));
    return set_scales_impl(plan, lm_y_scale, L_scale, U_scale);
}

int fdmgpu_set_refine(fdmgpu_plan plan, double rel_tol, int max_iter) {
    if (!plan || !(rel_tol > 0) || max_iter < 1 || max_iter > 64) return FDMGPU_ERR_INVALID_ARG;
    plan->rtol = rel_tol;
    plan->max_iter = max_iter;
    return FDMGPU_OK;
}

int fdmgpu_solve_device(fdmgpu_plan plan, void* d_ans, const void* d_rhs, void* stream) {
    if (!plan || !d_ans || !d_rhs) return FDMGPU_ERR_INVALID_ARG;
    return solve_device_impl(plan, d_ans, nullptr, d_rhs, nullptr, 1, (hipStream_t)stream);
}

int fdmgpu_solve_batch_device(fdmgpu_plan plan, int count, void* d_ans, const void* d_rhs, void* stream) {
    if (!plan || count < 0 || (count > 0 && (!d_ans || !d_rhs))) return FDMGPU_ERR_INVALID_ARG;
    if (count > 65535) return FDMGPU_ERR_UNSUPPORTED;
    return solve_device_impl(plan, d_ans, nullptr, d_rhs, nullptr, count, (hipStream_t)stream);
}

int fdmgpu_solve_table_device(fdmgpu_plan plan, int count, void* const* d_ans_tab, const void* const* d_rhs_tab,
                              void* stream) {
    if (!plan || count < 0 || (count > 0 && (!d_ans_tab || !d_rhs_tab))) return FDMGPU_ERR_INVALID_ARG;
    if (count > 65535) return FDMGPU_ERR_UNSUPPORTED;
    return solve_device_impl(plan, nullptr, (Tab)d_ans_tab, nullptr, d_rhs_tab, count, (hipStream_t)stream);
}

int fdmgpu_plan_variants(fdmgpu_plan plan, int max, fdmgpu_variant* out, int* count) {
    if (!plan || max < 0 || (max > 0 && !out)) return FDMGPU_ERR_INVALID_ARG;
    const std::vector<Variant> v = plan_variants(plan);
    for (int i = 0; i < max && i < (int)v.size(); i++)
        out[i] = {v[i].plane, v[i].tile_x, v[i].tile_y, v[i].zcr_gmul, v[i].fuse};
    if (count) *count = (int)v.size();
    return FDMGPU_OK;
}

int fdmgpu_plan_set_variant(fdmgpu_plan plan, const fdmgpu_variant* v) {
    if (!plan || !v) return FDMGPU_ERR_INVALID_ARG;
    const Variant w{v->plane, v->tile_x, v->tile_y, v->zcr_gmul, v->fuse};
    if (!variant_valid(plan, w)) return FDMGPU_ERR_INVALID_ARG;
    plan->var = w;
    return FDMGPU_OK;
}

int fdmgpu_plan_get_variant(fdmgpu_plan plan, fdmgpu_variant* v) {
    if (!plan || !v) return FDMGPU_ERR_INVALID_ARG;
    *v = {plan->var.plane, plan->var.tile_x, plan->var.tile_y, plan->var.zcr_gmul, plan->var.fuse};
    return FDMGPU_OK;
}

int fdmgpu_solve_host(fdmgpu_plan plan, void* ans, const void* rhs) {
    if (!plan || !ans || !rhs) return FDMGPU_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(plan->device));
    const size_t bytes = (size_t)plan->nx * plan->ny * plan->nz * plan->elem;
    const size_t need = plan->precision == FDMGPU_F32_REFINE64 ? 2 * bytes : bytes;
    if (plan->stage_bytes < need) {
        hipFree(plan->stage);
        plan->stage = nullptr;
        plan->stage_bytes = 0;
        HIPCHK(hipMalloc(&plan->stage, need));
        plan->stage_bytes = need;
    }
    if (!plan->own) HIPCHK(hipStreamCreateWithFlags(&plan->own, hipStreamNonBlocking));
    char* d_rhs = (char*)plan->stage;
    char* d_ans = plan->precision == FDMGPU_F32_REFINE64 ? d_rhs + bytes : d_rhs;
    HIPCHK(hipMemcpyAsync(d_rhs, rhs, bytes, hipMemcpyHostToDevice, plan->own));
    int rc = solve_device_impl(plan, d_ans, nullptr, d_rhs, nullptr, 1, plan->own);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(ans, d_ans, bytes, hipMemcpyDeviceToHost, plan->own));
    HIPCHK(hipStreamSynchronize(plan->own));
    if (plan->precision == FDMGPU_F32_REFINE64) {
        fdmgpu_stats st;
        rc = fdmgpu_get_stats(plan, &st);
        if (rc) return rc;
        if (!(st.rel_residual <= plan->rtol)) return FDMGPU_ERR_NOT_CONVERGED;
    }
    return FDMGPU_OK;
}

int fdmgpu_get_stats(fdmgpu_plan plan, fdmgpu_stats* st) {
    if (!plan || !st) return FDMGPU_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(plan->device));
    HIPCHK(hipStreamSynchronize(plan->last_stream));
    st->iterations = 1;
    st->rel_residual = -1.0;
    st->max_rel_residual = -1.0;
    if (plan->precision == FDMGPU_F32_REFINE64 && plan->last_count > 0) {
        const int c = plan->last_count;
        std::vector<int> it(c);
        std::vector<double> r(c);
        const int* d_it = plan->last_iters ? plan->last_iters : plan->w.iters;
        const double* d_r = plan->last_resid ? plan->last_resid : plan->w.resid;
        if (!d_it || !d_r) return FDMGPU_ERR_INVALID_ARG;
        HIPCHK(hipMemcpy(it.data(), d_it, sizeof(int) * c, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(r.data(), d_r, sizeof(double) * c, hipMemcpyDeviceToHost));
        st->iterations = *std::max_element(it.begin(), it.end());
        st->rel_residual = r[c - 1];
        st->max_rel_residual = *std::max_element(r.begin(), r.end());
    }
    return FDMGPU_OK;
}

int fdmgpu_plan_footprint(fdmgpu_plan plan, int count, unsigned long long* resident,
                          unsigned long long* per_solve) {
    if (!plan || count < 0) return FDMGPU_ERR_INVALID_ARG;
    const unsigned long long tabs = 3ull * ((1ull << plan->logx) + (1ull << plan->logy)) * plan->csize +
                                    (unsigned long long)(plan->nx + plan->ny) * plan->csize;
    if (resident) *resident = tabs;
    if (per_solve) {
        *per_solve = work_bytes(plan, count);
    }
    return FDMGPU_OK;
}

int fdmgpu_set_timing_kernels(fdmgpu_plan plan, unsigned mask) {
    if (!plan) return FDMGPU_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(plan->device));
    int rc = harvest_timing(plan);
    if (rc) return rc;
    plan->timing = mask & ((1u << FDMGPU_NUM_KERNELS) - 1);
    for (int k = 0; k < FDMGPU_NUM_KERNELS; k++) {
        plan->tot_ms[k] = 0;
        plan->launches[k] = 0;
    }
    return FDMGPU_OK;
}

int fdmgpu_set_timing(fdmgpu_plan plan, int enable) {
    return fdmgpu_set_timing_kernels(plan, enable ? ~0u : 0u);
}

int fdmgpu_get_timing(fdmgpu_plan plan, int kernel, double* total_ms, int* launches) {
    if (!plan || kernel < 0 || kernel >= FDMGPU_NUM_KERNELS) return FDMGPU_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(plan->device));
    int rc = harvest_timing(plan);
    if (rc) return rc;
    if (total_ms) *total_ms = plan->tot_ms[kernel];
    if (launches) *launches = plan->launches[kernel];
    return FDMGPU_OK;
}

int fdmgpu_destroy(fdmgpu_plan plan) {
    if (!plan) return FDMGPU_ERR_INVALID_ARG;
    (void)hipSetDevice(plan->device);
    // the stream of the last solve may be gone already: wait for the device
    // instead, so no kernel still reads the tables or scratch freed below
    (void)hipDeviceSynchronize();
    free_plan(plan);
    return FDMGPU_OK;
}

}  // extern "C"
