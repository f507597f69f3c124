// Axis-tile kernels (gfx950): the transform along x of a tile of rows, or
// along y of a tile of columns of one plane (planes that do not fit LDS, or
// non-square planes).  The transform is the DST-I on Dirichlet axes and the
// half-complex real DFT pair on periodic ones (template parameter KIND).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "dst_device.hpp"
#include "kernel_util.hpp"
#include "kernels.hpp"

namespace fdmgpu {

// ---------------------------------------------------------------------------
// axis-tile kernels (planes that do not fit LDS, or non-square planes)
// ---------------------------------------------------------------------------
// TILE = pencils per block (16, 32 or 64; a tunable, see tile_ok()).
template <int LOGN, typename T, int TILE_>
struct TileCfg {
    using S = DstShape<LOGN>;
    static constexpr int N = S::N;
    static constexpr int TILE = TILE_;
    static constexpr int SWN1 = (TILE / 2 > RowSwizzle<S::N1>::FMAX) ? S::N1 : 1;   // swizzle only if it fits
    static constexpr int P = N + S::PAD;            // x tiles: pencil pitch
    static constexpr int SJ = TILE + S::N1;         // y tiles: slot pitch
    static constexpr int WORK = (TILE / 2) * S::N1;
    static constexpr int NT = WORK < 64 ? 64 : WORK;
    static constexpr size_t lds_x() { return sizeof(T) * (size_t)TILE * P; }
    static constexpr size_t lds_y() { return sizeof(T) * (size_t)N * SJ; }
};

// rows = nrows pencils of IO::n points (N-1 Dirichlet, N periodic), contiguous;
// batch via blockIdx.y.
template <typename Tin, typename Tout, typename T, int LOGN, bool ACCUM, int KIND, int TILE_>
__global__ __launch_bounds__((TileCfg<LOGN, T, TILE_>::NT)) void xaxis_dst_kernel(
    const BatchPtr<const Tin> in, const BatchPtr<Tout> out, const long nrows, const T scale,
    const T* __restrict__ tables, const int* __restrict__ done) {
    using C = TileCfg<LOGN, T, TILE_>;
    using S = DstShape<LOGN>;
    using IO = PencilIO<KIND, LOGN>;
    using LX = PlaneLayout<C::P, S::N1, true>;
    if (done && done[blockIdx.y]) return;
    constexpr int n = IO::n, o = IO::off, NT = C::NT, TILE = C::TILE;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* lds = reinterpret_cast<T*>(smem_raw);
    const int tid = threadIdx.x;
    const long r0 = (long)blockIdx.x * TILE;
    const auto src = as_global(in(blockIdx.y) + r0 * n);
    const auto dst = as_global(out(blockIdx.y) + r0 * n);
    const int rows_here = (int)((nrows - r0) < TILE ? (nrows - r0) : TILE);

    DstLane<T, LOGN> L;
    L.init(tables, tid & (S::N1 - 1));
    if constexpr (KIND == kDst)
        for (int p = tid; p < TILE; p += NT) lds[LX::at(p, 0)] = T(0);
    for (int idx = tid; idx < TILE * n; idx += NT) {
        const int p = idx / n, x = idx - p * n;
        lds[LX::at(p, x + o)] = p < rows_here ? (T)src[idx] : T(0);
    }
    __syncthreads();
    pencil_pairs<KIND, T, LOGN, TILE / 2, LX, NT>(lds, L, scale, tid);
    __syncthreads();
    for (int idx = tid; idx < rows_here * n; idx += NT) {
        const int p = idx / n, x = idx - p * n;
        const T v = lds[LX::at(p, IO::out(x))];
        if constexpr (ACCUM) dst[idx] += (Tout)v;
        else dst[idx] = (Tout)v;
    }
}

// Columns of plane blockIdx.y (nplanes planes of [IO::n][nx]); batch via blockIdx.z.
template <typename Tin, typename Tout, typename T, int LOGN, int KIND, int TILE_>
__global__ __launch_bounds__((TileCfg<LOGN, T, TILE_>::NT)) void yaxis_dst_kernel(
    const BatchPtr<const Tin> in, const BatchPtr<Tout> out, const int nx, const T scale,
    const T* __restrict__ tables, const int* __restrict__ done) {
    using C = TileCfg<LOGN, T, TILE_>;
    using S = DstShape<LOGN>;
    using IO = PencilIO<KIND, LOGN>;
    using LY = PlaneLayout<C::SJ, C::SWN1, false>;   // rows = slots, columns = pencils
    constexpr int n = IO::n, o = IO::off, NT = C::NT, TILE = C::TILE;
    if (done && done[blockIdx.z]) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* lds = reinterpret_cast<T*>(smem_raw);
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * TILE;
    const long poff = (long)blockIdx.y * n * nx;
    const auto src = as_global(in(blockIdx.z) + poff);
    const auto dst = as_global(out(blockIdx.z) + poff);

    DstLane<T, LOGN> L;
    L.init(tables, tid & (S::N1 - 1));
    if constexpr (KIND == kDst)
        for (int p = tid; p < TILE; p += NT) lds[LY::addr(p, 0)] = T(0);   // slot 0 of every pencil
    for (int idx = tid; idx < n * TILE; idx += NT) {
        const int y = idx / TILE, p = idx - y * TILE;
        const int x = x0 + p;
        lds[LY::addr(p, y + o)] = x < nx ? (T)src[(long)y * nx + x] : T(0);
    }
    __syncthreads();
    pencil_pairs<KIND, T, LOGN, TILE / 2, LY, NT>(lds, L, scale, tid);
    __syncthreads();
    for (int idx = tid; idx < n * TILE; idx += NT) {
        const int y = idx / TILE, p = idx - y * TILE;
        const int x = x0 + p;
        if (x < nx) dst[(long)y * nx + x] = (Tout)lds[LY::addr(p, IO::out(y))];
    }
}


// Tile sizes an axis kernel of length 2^logn supports: 16 always; 32 and 64
// for N >= 32 when the y tile (N x (TILE + N1) elements) fits LDS.
template <typename T>
bool tile_ok(int logn, int tile) {
    const size_t N = (size_t)1 << logn, N1 = (size_t)1 << (logn / 2);
    if (tile == 16) return true;
    if (tile != 32 && tile != 64) return false;
    return logn >= 5 && N * (tile + N1) * sizeof(T) <= kLdsBytes;
}

// 32 pencils per block, or 16 when a 32-wide y tile would not fit LDS
template <typename T>
int default_tile(int logn) {
    return tile_ok<T>(logn, 32) ? 32 : 16;
}

template <typename Tin, typename Tout, typename T, int LOGN, bool ACCUM, int KIND, int TILE>
static hipError_t xaxis_k(const AxisArgs& a) {
    using C = TileCfg<LOGN, T, TILE>;
    auto k = xaxis_dst_kernel<Tin, Tout, T, LOGN, ACCUM, KIND, TILE>;
    const size_t lds = C::lds_x();
    hipError_t e = allow_lds(k, lds);
    if (e != hipSuccess) return e;
    const long tiles = (a.nrows + C::TILE - 1) / C::TILE;
    const BatchPtr<const Tin> in{(const Tin*)a.in, a.in_tab, a.bs_in};
    const BatchPtr<Tout> out{(Tout*)a.out, a.out_tab, a.bs_out};
    hipLaunchKernelGGL(k, dim3((unsigned)tiles, a.batch), dim3(C::NT), lds, a.stream, in, out, a.nrows,
                       (T)a.scale, (const T*)a.tables, a.done);
    return hipGetLastError();
}

template <typename Tin, typename Tout, typename T, int LOGN, int TILE>
static hipError_t xaxis_tile(const AxisArgs& a) {
    if (a.kind == kDst)
        return a.accum ? xaxis_k<Tin, Tout, T, LOGN, true, kDst, TILE>(a)
                       : xaxis_k<Tin, Tout, T, LOGN, false, kDst, TILE>(a);
    if constexpr (std::is_same_v<Tin, T> && std::is_same_v<Tout, T>) {
        if (a.accum) return hipErrorInvalidValue;
        if (a.kind == kRdftFwd) return xaxis_k<Tin, Tout, T, LOGN, false, kRdftFwd, TILE>(a);
        if (a.kind == kRdftInv) return xaxis_k<Tin, Tout, T, LOGN, false, kRdftInv, TILE>(a);
    }
    return hipErrorInvalidValue;
}

template <typename Tin, typename Tout, typename T, int LOGN>
static hipError_t xaxis_t(const AxisArgs& a) {
    const int tile = a.tile ? a.tile : default_tile<T>(LOGN);
    if (!tile_ok<T>(LOGN, tile)) return hipErrorInvalidValue;
    if (tile == 16) return xaxis_tile<Tin, Tout, T, LOGN, 16>(a);
    if constexpr (LOGN >= 5) {
        if (tile == 32) return xaxis_tile<Tin, Tout, T, LOGN, 32>(a);
        if constexpr ((size_t)(1 << LOGN) * (64 + (1 << (LOGN / 2))) * sizeof(T) <= kLdsBytes)
            if (tile == 64) return xaxis_tile<Tin, Tout, T, LOGN, 64>(a);
    }
    return hipErrorInvalidValue;
}

template <typename Tin, typename Tout, typename T, int LOGN, int KIND, int TILE>
static hipError_t yaxis_k(const AxisArgs& a) {
    using C = TileCfg<LOGN, T, TILE>;
    auto k = yaxis_dst_kernel<Tin, Tout, T, LOGN, KIND, TILE>;
    const size_t lds = C::lds_y();
    hipError_t e = allow_lds(k, lds);
    if (e != hipSuccess) return e;
    const int tiles = (a.nx + C::TILE - 1) / C::TILE;
    const BatchPtr<const Tin> in{(const Tin*)a.in, a.in_tab, a.bs_in};
    const BatchPtr<Tout> out{(Tout*)a.out, a.out_tab, a.bs_out};
    hipLaunchKernelGGL(k, dim3(tiles, a.nplanes, a.batch), dim3(C::NT), lds, a.stream, in, out, a.nx,
                       (T)a.scale, (const T*)a.tables, a.done);
    return hipGetLastError();
}

template <typename Tin, typename Tout, typename T, int LOGN, int TILE>
static hipError_t yaxis_tile(const AxisArgs& a) {
    if (a.kind == kDst) return yaxis_k<Tin, Tout, T, LOGN, kDst, TILE>(a);
    if constexpr (std::is_same_v<Tin, T> && std::is_same_v<Tout, T>) {
        if (a.kind == kRdftFwd) return yaxis_k<Tin, Tout, T, LOGN, kRdftFwd, TILE>(a);
        if (a.kind == kRdftInv) return yaxis_k<Tin, Tout, T, LOGN, kRdftInv, TILE>(a);
    }
    return hipErrorInvalidValue;
}

template <typename Tin, typename Tout, typename T, int LOGN>
static hipError_t yaxis_t(const AxisArgs& a) {
    const int tile = a.tile ? a.tile : default_tile<T>(LOGN);
    if (!tile_ok<T>(LOGN, tile)) return hipErrorInvalidValue;
    if (tile == 16) return yaxis_tile<Tin, Tout, T, LOGN, 16>(a);
    if constexpr (LOGN >= 5) {
        if (tile == 32) return yaxis_tile<Tin, Tout, T, LOGN, 32>(a);
        if constexpr ((size_t)(1 << LOGN) * (64 + (1 << (LOGN / 2))) * sizeof(T) <= kLdsBytes)
            if (tile == 64) return yaxis_tile<Tin, Tout, T, LOGN, 64>(a);
    }
    return hipErrorInvalidValue;
}

#define FDM_AXIS_SWITCH(FN)                      \
    switch (logn) {                              \
        case 1: return FN<Tin, Tout, T, 1>(a);   \
        case 2: return FN<Tin, Tout, T, 2>(a);   \
        case 3: return FN<Tin, Tout, T, 3>(a);   \
        case 4: return FN<Tin, Tout, T, 4>(a);   \
        case 5: return FN<Tin, Tout, T, 5>(a);   \
        case 6: return FN<Tin, Tout, T, 6>(a);   \
        case 7: return FN<Tin, Tout, T, 7>(a);   \
        case 8: return FN<Tin, Tout, T, 8>(a);   \
        case 9: return FN<Tin, Tout, T, 9>(a);   \
        default: return hipErrorInvalidValue;    \
    }

template <typename Tin, typename Tout, typename T>
hipError_t launch_xaxis(int logn, const AxisArgs& a) {
    FDM_AXIS_SWITCH(xaxis_t)
}
template <typename Tin, typename Tout, typename T>
hipError_t launch_yaxis(int logn, const AxisArgs& a) {
    FDM_AXIS_SWITCH(yaxis_t)
}

template hipError_t launch_xaxis<float, float, float>(int, const AxisArgs&);
template hipError_t launch_xaxis<double, double, double>(int, const AxisArgs&);
template hipError_t launch_xaxis<double, float, float>(int, const AxisArgs&);
template hipError_t launch_xaxis<float, double, float>(int, const AxisArgs&);
template hipError_t launch_yaxis<float, float, float>(int, const AxisArgs&);
template hipError_t launch_yaxis<double, double, double>(int, const AxisArgs&);
template hipError_t launch_yaxis<double, float, float>(int, const AxisArgs&);
template hipError_t launch_yaxis<float, double, float>(int, const AxisArgs&);
template bool tile_ok<float>(int, int);
template bool tile_ok<double>(int, int);
template int default_tile<float>(int);
template int default_tile<double>(int);

}  // namespace fdmgpu
