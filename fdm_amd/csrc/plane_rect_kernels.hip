// Plane kernels for rectangular planes (gfx950): the DST-I along x and y of
// whole z-planes of ny x nx points, nx + 1 != ny + 1 (both powers of two),
// held in LDS -- the same transforms as the square plane kernel, with a
// transform shape per axis.  A LaplCube whose planes are not square (e.g.
// 127 x 63 x 31) thus takes three passes over the field per solve (xy
// transforms, z cyclic reduction, inverse xy transforms) instead of the five
// of the axis-tile path (x, y, z, y, x).
//
// Layout in HBM: dense [nz][ny][nx] (x fastest) per problem, problems back to
// back or through a pointer table (reference src/lapl_cube.h:80).  In LDS the
// plane is NY rows (row 0 the DST's dummy) of pitch NX + N1x slots, column 0
// the dummy; the row swizzle period is the smaller of the two axes' lanes per
// job, so the column pass's compile-time slot offsets stay swizzle-invariant.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "dst_device.hpp"
#include "kernel_util.hpp"
#include "kernels.hpp"

namespace fdmgpu {

template <int LXN, int LYN>
struct RectCfg {
    using SX = DstShape<LXN>;
    using SY = DstShape<LYN>;
    static constexpr int NX = SX::N, NY = SY::N;   // transform lengths: nx + 1, ny + 1
    static constexpr int N1L = SX::N1 < SY::N1 ? SX::N1 : SY::N1;
    static constexpr int P = NX + SX::N1;          // pitch = N1x mod 32: conflict-free row passes
    static constexpr int WX = (NY / 2) * SX::N1, WY = (NX / 2) * SY::N1;
    static constexpr int W = WX > WY ? WX : WY;
    static constexpr int NT = W < 64 ? 64 : W;
    static_assert(NT <= 1024 && NT % NX == 0, "block covers whole rows");
    template <typename T>
    static constexpr size_t lds_bytes() {
        return sizeof(T) * (size_t)NY * P;
    }
    // waves per SIMD: what LDS allows, and at most 2 for an FP64 transform of
    // length 128, whose lane constants and values need about 190 VGPRs
    template <typename T>
    static constexpr int waves() {
        const int w = lds_waves_per_simd(lds_bytes<T>(), NT);
        return sizeof(T) == 8 && (LXN == 7 || LYN == 7) && w > 2 ? 2 : w;
    }
};

// ACC: 0 out = result; 1 out += result; 2 out = base + result with an FP32
// base (the refinement's first iterate), and out = base for problems flagged
// done.
template <typename Tin, typename Tout, typename T, int LXN, int LYN, int ACC>
__global__ __launch_bounds__((RectCfg<LXN, LYN>::NT))
__attribute__((amdgpu_waves_per_eu(RectCfg<LXN, LYN>::template waves<T>()))) void plane_rect_kernel(
    const BatchPtr<const Tin> in, const BatchPtr<Tout> out, const T sx, const T sy, const T* __restrict__ tab_x,
    const T* __restrict__ tab_y, const int* __restrict__ done, const float* __restrict__ base, const long base_bs) {
    using C = RectCfg<LXN, LYN>;
    using SX = typename C::SX;
    using SY = typename C::SY;
    using LayX = PlaneLayout<C::P, C::N1L, true>;    // pencils = rows
    using LayY = PlaneLayout<C::P, C::N1L, false>;   // pencils = columns
    constexpr int NX = C::NX, NY = C::NY, NT = C::NT;
    constexpr int nx = NX - 1, ny = NY - 1, R = NT / NX;
    constexpr int ITER = (ny + R - 1) / R;
    const int tid = threadIdx.x;
    const int x = tid & (NX - 1), r0 = tid / NX;
    const long plane_off = (long)blockIdx.x * nx * ny;
    const auto dst = as_global(out(blockIdx.y) + plane_off);
    [[maybe_unused]] const auto bp = as_global(ACC == 2 ? base + blockIdx.y * base_bs + plane_off : base);
    // the sweep over the plane's points: thread (x, r0) takes rows r0 + R i
    auto sweep = [&](auto&& f) {
        if (x < nx) {
#pragma unroll
            for (int i = 0; i < ITER; ++i) {
                const int y = r0 + R * i;
                if (R * i + R <= ny || y < ny) f(i, y);
            }
        }
    };
    if (done && done[blockIdx.y]) {
        if constexpr (ACC == 2) sweep([&](int, int y) { dst[y * nx + x] = (Tout)bp[y * nx + x]; });
        return;
    }
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* lds = reinterpret_cast<T*>(smem_raw);
    for (int i = tid; i < NX; i += NT) lds[LayX::at(0, i)] = T(0);   // dummy row 0
    for (int i = tid; i < NY; i += NT) lds[LayX::at(i, 0)] = T(0);   // dummy column 0
    {   // stage: every global load in flight before the first LDS store
        const auto src = as_global(in(blockIdx.y) + plane_off);
        Tin v[ITER];
        sweep([&](int i, int y) { v[i] = src[y * nx + x]; });
        __builtin_amdgcn_sched_barrier(0);
        sweep([&](int i, int y) { lds[LayX::at(y + 1, x + 1)] = (T)v[i]; });
    }
    __syncthreads();
    {
        DstLane<T, LXN> L;
        L.init(tab_x, tid & (SX::N1 - 1));
        pencil_pairs<kDst, T, LXN, NY / 2, LayX, NT>(lds, L, sx, tid);
    }
    __syncthreads();
    {
        int ty = tid;   // the y lane's constants are loaded after the x pass
        asm volatile("" : "+v"(ty));
        DstLane<T, LYN> L;
        L.init(tab_y, ty & (SY::N1 - 1));
        pencil_pairs<kDst, T, LYN, NX / 2, LayY, NT>(lds, L, sy, ty);
    }
    __syncthreads();
    // output point (y, x) sits in row sigma_y(y + 1), column sigma_x(x + 1)
    if constexpr (ACC == 2) {
        float b[ITER];
        sweep([&](int i, int y) { b[i] = bp[y * nx + x]; });
        __builtin_amdgcn_sched_barrier(0);
        sweep([&](int i, int y) {
            dst[y * nx + x] = (Tout)b[i] + (Tout)lds[LayX::at(SY::sigma(y + 1), SX::sigma(x + 1))];
        });
    } else {
        sweep([&](int, int y) {
            const T v = lds[LayX::at(SY::sigma(y + 1), SX::sigma(x + 1))];
            if constexpr (ACC == 1) dst[y * nx + x] += (Tout)v;
            else dst[y * nx + x] = (Tout)v;
        });
    }
}

template <typename Tin, typename Tout, typename T, int LXN, int LYN, int ACC>
static hipError_t rect_k(const PlaneArgs& a) {
    using C = RectCfg<LXN, LYN>;
    auto k = plane_rect_kernel<Tin, Tout, T, LXN, LYN, ACC>;
    const size_t lds = C::template lds_bytes<T>();
    hipError_t e = allow_lds(k, lds);
    if (e != hipSuccess) return e;
    const BatchPtr<const Tin> in{(const Tin*)a.in, a.in_tab, a.bs_in};
    const BatchPtr<Tout> out{(Tout*)a.out, a.out_tab, a.bs_out};
    hipLaunchKernelGGL(k, dim3(a.nplanes, a.batch), dim3(C::NT), lds, a.stream, in, out, (T)a.sx, (T)a.sy,
                       (const T*)a.tables, (const T*)a.tables_y, a.done, a.base, a.bs_out);
    return hipGetLastError();
}

template <typename Tin, typename Tout, typename T, int LXN, int LYN>
static hipError_t rect_t(const PlaneArgs& a) {
    if (a.kind != kDst || a.stream_planes) return hipErrorInvalidValue;
    if (a.base && !a.accum) return hipErrorInvalidValue;
    if constexpr (std::is_same_v<Tout, double> && !std::is_same_v<Tin, double>) {
        if (a.base) return rect_k<Tin, Tout, T, LXN, LYN, 2>(a);
        if (a.accum) return rect_k<Tin, Tout, T, LXN, LYN, 1>(a);
    }
    if (a.base || a.accum) return hipErrorInvalidValue;
    return rect_k<Tin, Tout, T, LXN, LYN, 0>(a);
}

// transform lengths 2^4 .. 2^7 per axis (interior 15 .. 127), unequal
constexpr int kRectMin = 4, kRectMax = 7;

template <typename Tin, typename Tout, typename T, int LXN, int LYN>
static hipError_t rect_if(const PlaneArgs& a) {
    if constexpr (LXN != LYN) return rect_t<Tin, Tout, T, LXN, LYN>(a);
    else return hipErrorInvalidValue;   // square planes: the plane kernel
}

template <typename Tin, typename Tout, typename T, int LXN>
static hipError_t rect_y(int logy, const PlaneArgs& a) {
    switch (logy) {
        case 4: return rect_if<Tin, Tout, T, LXN, 4>(a);
        case 5: return rect_if<Tin, Tout, T, LXN, 5>(a);
        case 6: return rect_if<Tin, Tout, T, LXN, 6>(a);
        case 7: return rect_if<Tin, Tout, T, LXN, 7>(a);
        default: return hipErrorInvalidValue;
    }
}

template <typename Tin, typename Tout, typename T>
hipError_t launch_plane_rect(int logx, int logy, const PlaneArgs& a) {
    switch (logx) {
        case 4: return rect_y<Tin, Tout, T, 4>(logy, a);
        case 5: return rect_y<Tin, Tout, T, 5>(logy, a);
        case 6: return rect_y<Tin, Tout, T, 6>(logy, a);
        case 7: return rect_y<Tin, Tout, T, 7>(logy, a);
        default: return hipErrorInvalidValue;
    }
}

bool plane_rect_fits(int logx, int logy, bool f64) {
    if (logx == logy || logx < kRectMin || logx > kRectMax || logy < kRectMin || logy > kRectMax) return false;
    const size_t nxp = (size_t)1 << logx, ny = (size_t)1 << logy, n1x = (size_t)1 << (logx / 2);
    return (f64 ? 8 : 4) * ny * (nxp + n1x) <= kLdsBytes;
}

template hipError_t launch_plane_rect<float, float, float>(int, int, const PlaneArgs&);
template hipError_t launch_plane_rect<double, double, double>(int, int, const PlaneArgs&);
template hipError_t launch_plane_rect<double, float, float>(int, int, const PlaneArgs&);
template hipError_t launch_plane_rect<float, double, float>(int, int, const PlaneArgs&);

}  // namespace fdmgpu
