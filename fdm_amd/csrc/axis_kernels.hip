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
// TILE = pencils per block (16, 32 or 64; a tunable, see tile_ok()).  Both
// kernels hold their tile in LDS as TILE pencils of pitch P = N + N1 (the
// pencils are rows, the layout of the plane kernels' x pass), so the
// transform's LDS accesses are the same for x and y tiles.
template <int LOGN, typename T, int TILE_>
struct TileCfg {
    using S = DstShape<LOGN>;
    static constexpr int N = S::N;
    static constexpr int TILE = TILE_;
    static constexpr int P = N + S::PAD;   // pencil pitch
    static constexpr int WORK = (TILE / 2) * S::N1;
    static constexpr int NT = WORK < 64 ? 64 : WORK;
    static constexpr size_t lds() { return sizeof(T) * (size_t)TILE * P; }
};

constexpr size_t tile_lds_bytes(int logn, int tile, size_t elem) {
    return elem * (size_t)tile * (((size_t)1 << logn) + ((size_t)1 << (logn / 2)));
}

// Global -> LDS copy of ITEMS items over NT threads (thread tid takes items
// tid, tid + NT, ...), in batches of KB: every load of a batch is issued into
// registers before the batch's first LDS store, so up to KB loads per thread
// are in flight at once.  ld(idx) is called for idx < ITER * NT and must stay
// in bounds for all of them (the callers clamp the address and select zero);
// st(idx, v) is called for idx < ITEMS only.
template <int ITEMS, int NT, int KB, typename V, typename LD, typename ST>
__device__ __forceinline__ void staged_copy(const int tid, LD&& ld, ST&& st) {
    constexpr int ITER = (ITEMS + NT - 1) / NT;
#pragma unroll
    for (int b0 = 0; b0 < ITER; b0 += KB) {
        V v[KB];
#pragma unroll
        for (int i = 0; i < KB; ++i)
            if (b0 + i < ITER) v[i] = ld(tid + NT * (b0 + i));
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < KB; ++i) {
            const int idx = tid + NT * (b0 + i);
            if (b0 + i < ITER && (ITEMS % NT == 0 || idx < ITEMS)) st(idx, v[i]);
        }
    }
}

// Loads per thread in flight in the tile copy: 32 for FP64 transforms (the
// deeper batch measured faster at 255^3 than the occupancy a smaller one
// buys), 16 for FP32 ones (keeps them at 4 waves per SIMD).
template <typename T>
constexpr int kCopyBatch = sizeof(T) == 8 ? 32 : 16;

// Element map of an x tile (TILE rows of N slots): thread tid owns columns
// x = tid % XW + XW*c (c < CX) of rows r0 + R*i (i < RI), r0 = tid / XW,
// XW = min(NT, N).  Global element (row, x) of the tile is at row*n + x, so
// every element is a lane offset plus a compile-time one; the LDS rows of a
// thread repeat their swizzle f() every PER elements (period N1 in the row).
template <int LOGN, int NT, int TILE, int P, int N1>
struct XTileMap {
    static constexpr int N = 1 << LOGN;
    static constexpr int XW = NT < N ? NT : N, CX = N / XW, R = NT / XW, RI = R < TILE ? TILE / R : 1;
    static constexpr int PER = N1 / gcd_int(R, N1);
    static_assert((R < TILE ? RI * R == TILE : true) && CX * XW == N, "tile map");
    using Lay = PlaneLayout<P, N1, true>;
    int xb, r0;
    bool on;   // the thread has rows (false only when R > TILE)
    __device__ explicit XTileMap(const int tid) : xb(tid % XW), r0(tid / XW), on(R <= TILE || tid / XW < TILE) {}
    // LDS row base of element i (column 0, f included): PER bases
    __device__ int row_base(const int j) const { return Lay::at(r0 + R * j, 0); }
};

// LDS <-> global sweep of an x tile in batches of KB elements per thread: the
// batch's loads are issued before its first store.  ld(c, i) / st(c, i, v)
// get element indices that are constants once the loops are unrolled;
// columns x >= NN are skipped, and so is everything when `on` is false.
template <int CX, int RI, int XW, int NN, int KB, typename V, typename LD, typename ST>
__device__ __forceinline__ void xtile_sweep(const int xb, const bool on, LD&& ld, ST&& st) {
#pragma unroll
    for (int c = 0; c < CX; ++c) {
        if (!on || xb + XW * c >= NN) continue;   // one branch per column, none per element
#pragma unroll
        for (int b0 = 0; b0 < RI; b0 += KB) {
            V v[KB];
#pragma unroll
            for (int e = 0; e < KB; ++e)
                if (b0 + e < RI) v[e] = ld(c, b0 + e);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int e = 0; e < KB; ++e)
                if (b0 + e < RI) st(c, b0 + e, v[e]);
        }
    }
}

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
    const int rows_here = (int)((nrows - r0) < TILE ? (nrows - r0) : TILE);
    // the tile's rows through buffer resources sized to the rows present:
    // rows past the last one read 0 and their stores are dropped
    const unsigned pts = (unsigned)rows_here * n;
    const BufRange<Tin> src(in(blockIdx.y) + r0 * n, pts * sizeof(Tin));
    const BufRange<Tout> dst(out(blockIdx.y) + r0 * n, pts * sizeof(Tout));
    using M = XTileMap<LOGN, NT, TILE, C::P, S::N1>;
    const M m(tid);
    int rb[M::PER];
#pragma unroll
    for (int j = 0; j < M::PER; ++j) rb[j] = opaque(m.row_base(j));
    // LDS slot of element (c, i) at column col: rows R*(i - i % PER) apart
    // share the swizzle
    auto lslot = [&](int i, int col) { return rb[i % M::PER] + M::R * (i - i % M::PER) * C::P + col; };
    const int lane_in = (m.r0 * n + m.xb) * (int)sizeof(Tin), lane_out = (m.r0 * n + m.xb) * (int)sizeof(Tout);

    DstLane<T, LOGN> L;
    L.init(tables, tid & (S::N1 - 1));
    if constexpr (KIND == kDst)
        for (int p = tid; p < TILE; p += NT) lds[LX::at(p, 0)] = T(0);
    xtile_sweep<M::CX, M::RI, M::XW, n, kCopyBatch<T>, Tin>(
        m.xb, m.on,
        [&](int c, int i) { return src.load(lane_in, (int)((M::XW * c + M::R * i * n) * sizeof(Tin))); },
        [&](int c, int i, Tin v) { lds[lslot(i, m.xb + M::XW * c + o)] = (T)v; });
    __syncthreads();
    pencil_pairs<KIND, T, LOGN, TILE / 2, LX, NT>(lds, L, scale, tid);
    __syncthreads();
    int oc[M::CX];
#pragma unroll
    for (int c = 0; c < M::CX; ++c) oc[c] = IO::out(m.xb + M::XW * c < n ? m.xb + M::XW * c : 0);
    if constexpr (ACCUM) {
        // out += result: a batch's global loads are all in flight before its
        // first add and store
        struct Acc {
            Tout old;
            T v;
        };
        xtile_sweep<M::CX, M::RI, M::XW, n, kCopyBatch<T>, Acc>(
            m.xb, m.on,
            [&](int c, int i) {
                const int so = (int)((M::XW * c + M::R * i * n) * sizeof(Tout));
                return Acc{dst.load(lane_out, so), lds[lslot(i, oc[c])]};
            },
            [&](int c, int i, Acc a) {
                const int so = (int)((M::XW * c + M::R * i * n) * sizeof(Tout));
                dst.store(a.old + (Tout)a.v, lane_out, so);
            });
    } else {
        xtile_sweep<M::CX, M::RI, M::XW, n, 2 * kCopyBatch<T>, T>(
            m.xb, m.on, [&](int c, int i) { return lds[lslot(i, oc[c])]; },
            [&](int c, int i, T v) {
                const int so = (int)((M::XW * c + M::R * i * n) * sizeof(Tout));
                dst.store((Tout)v, lane_out, so);
            });
    }
}

// Tiles of the y kernel dealt to the XCDs in contiguous runs: workgroups of a
// 1-D grid go to the 8 XCDs round-robin by blockIdx.x, and XCD x takes tiles
// [x per, (x+1) per) in (x tile, plane, problem) order.  A tile's row segments
// (TILE consecutive x of a row of pitch nx) share their end cache lines with
// the neighbouring tiles of the plane; with the tiles of a plane spread over
// all XCDs every XCD fetched those lines again (measured 1.3x the
// algorithmic bytes at 255^2 planes).  The grid is padded to a multiple of
// 8 blocks; blocks past the last tile return.
constexpr int kAxisXcds = 8;

// Columns of nplanes planes of [IO::n][nx] per problem, TILE columns per
// block.  Column p of the tile is stored as LDS row p: the global reads stay
// row segments of TILE consecutive x, and the transform is the x pass.
template <typename Tin, typename Tout, typename T, int LOGN, int KIND, int TILE_>
__global__ __launch_bounds__((TileCfg<LOGN, T, TILE_>::NT)) void yaxis_dst_kernel(
    const BatchPtr<const Tin> in, const BatchPtr<Tout> out, const int nx, const T scale,
    const T* __restrict__ tables, const int* __restrict__ done, const int nplanes, const int batch) {
    using C = TileCfg<LOGN, T, TILE_>;
    using S = DstShape<LOGN>;
    using IO = PencilIO<KIND, LOGN>;
    using LT = PlaneLayout<C::P, S::N1, true>;   // rows = pencils (columns of the plane)
    constexpr int n = IO::n, o = IO::off, NT = C::NT, TILE = C::TILE;
    const int ntx = (nx + TILE - 1) / TILE;
    const long per = (long)gridDim.x / kAxisXcds;
    const long tile_id = ((long)blockIdx.x % kAxisXcds) * per + (long)blockIdx.x / kAxisXcds;
    if (tile_id >= (long)ntx * nplanes * batch) return;
    const int bx = (int)(tile_id % ntx);
    const long rest = tile_id / ntx;
    const int by = (int)(rest % nplanes), bz = (int)(rest / nplanes);
    if (done && done[bz]) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* lds = reinterpret_cast<T*>(smem_raw);
    const int tid = threadIdx.x;
    const int x0 = bx * TILE;
    const long poff = (long)by * n * nx;
    // thread tid owns tile column p = tid % TILE (plane column x0 + p) and
    // rows y0 + YR*i, y0 = tid / TILE: the LDS slot of a row is the pencil's
    // base plus a compile-time offset, the global one a lane offset plus a
    // wave-uniform one (buffer resources over the plane)
    static_assert(NT % TILE == 0, "whole tile rows per pass");
    constexpr int YR = NT / TILE, YI = (n + YR - 1) / YR;
    const unsigned pts = (unsigned)n * nx;
    const BufRange<Tin> src(in(bz) + poff, pts * sizeof(Tin));
    const BufRange<Tout> dst(out(bz) + poff, pts * sizeof(Tout));
    const int p = tid % TILE, y0 = tid / TILE;
    const bool col_ok = x0 + p < nx;
    const int lane_in = (y0 * nx + (col_ok ? x0 + p : 0)) * (int)sizeof(Tin);
    const int lane_out = (y0 * nx + (col_ok ? x0 + p : 0)) * (int)sizeof(Tout);
    const int pb = opaque(LT::at(p, 0));

    DstLane<T, LOGN> L;
    L.init(tables, tid & (S::N1 - 1));
    if constexpr (KIND == kDst)
        for (int q = tid; q < TILE; q += NT) lds[LT::at(q, 0)] = T(0);   // slot 0 of every pencil
    constexpr int KB = kCopyBatch<T>;
#pragma unroll
    for (int b0 = 0; b0 < YI; b0 += KB) {
        Tin v[KB];
#pragma unroll
        for (int e = 0; e < KB; ++e) {
            const int i = b0 + e;
            if (i < YI && (YR * i + YR <= n || y0 + YR * i < n)) {
                const Tin raw = src.load(lane_in, YR * i * nx * (int)sizeof(Tin));
                v[e] = col_ok ? raw : Tin(0);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int e = 0; e < KB; ++e) {
            const int i = b0 + e;
            if (i < YI && (YR * i + YR <= n || y0 + YR * i < n)) lds[pb + y0 + YR * i + o] = (T)v[e];
        }
    }
    __syncthreads();
    pencil_pairs<KIND, T, LOGN, TILE / 2, LT, NT>(lds, L, scale, tid);
    __syncthreads();
    if (col_ok) {
#pragma unroll 16
        for (int i = 0; i < YI; ++i) {
            const int y = y0 + YR * i;
            if (YR * i + YR <= n || y < n)
                dst.store((Tout)lds[pb + IO::out(y)], lane_out, YR * i * nx * (int)sizeof(Tout));
        }
    }
}


// Tile sizes an axis kernel of length 2^logn supports: 16 always; 32 and 64
// for N >= 32 when the tile (TILE x (N + N1) elements) fits LDS.
template <typename T>
bool tile_ok(int logn, int tile) {
    if (tile == 16) return true;
    if (tile != 32 && tile != 64) return false;
    return logn >= 5 && tile_lds_bytes(logn, tile, sizeof(T)) <= kLdsBytes;
}

// 32 pencils per block, or 16 when a 32-pencil tile would not fit LDS
template <typename T>
int default_tile(int logn) {
    return tile_ok<T>(logn, 32) ? 32 : 16;
}

// y tiles of FP32 planes at least 128 wide: 64 pencils when two such blocks
// fit a CU (measured 4% faster than 32 for the 255^3 FP32+refine solve)
template <typename T>
int default_tile_y(int logn, int nx) {
    if (sizeof(T) == 4 && nx >= 128 && tile_ok<T>(logn, 64) && 2 * tile_lds_bytes(logn, 64, 4) <= kLdsBytes)
        return 64;
    return default_tile<T>(logn);
}

template <typename Tin, typename Tout, typename T, int LOGN, bool ACCUM, int KIND, int TILE>
static hipError_t xaxis_k(const AxisArgs& a) {
    using C = TileCfg<LOGN, T, TILE>;
    auto k = xaxis_dst_kernel<Tin, Tout, T, LOGN, ACCUM, KIND, TILE>;
    const size_t lds = C::lds();
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
        if constexpr (tile_lds_bytes(LOGN, 32, sizeof(T)) <= kLdsBytes)
            if (tile == 32) return xaxis_tile<Tin, Tout, T, LOGN, 32>(a);
        if constexpr (tile_lds_bytes(LOGN, 64, sizeof(T)) <= kLdsBytes)
            if (tile == 64) return xaxis_tile<Tin, Tout, T, LOGN, 64>(a);
    }
    return hipErrorInvalidValue;
}

template <typename Tin, typename Tout, typename T, int LOGN, int KIND, int TILE>
static hipError_t yaxis_k(const AxisArgs& a) {
    using C = TileCfg<LOGN, T, TILE>;
    auto k = yaxis_dst_kernel<Tin, Tout, T, LOGN, KIND, TILE>;
    // the kernel addresses a plane through 32-bit byte offsets
    constexpr size_t elem = sizeof(Tin) > sizeof(Tout) ? sizeof(Tin) : sizeof(Tout);
    if (((size_t)1 << LOGN) * (size_t)a.nx * elem >= ((size_t)1 << 31)) return hipErrorInvalidValue;
    const size_t lds = C::lds();
    hipError_t e = allow_lds(k, lds);
    if (e != hipSuccess) return e;
    const long tiles = (long)((a.nx + C::TILE - 1) / C::TILE) * a.nplanes * a.batch;
    const long blocks = (tiles + kAxisXcds - 1) / kAxisXcds * kAxisXcds;
    if (tiles < 1 || blocks > 0x7fffffffL) return hipErrorInvalidValue;
    const BatchPtr<const Tin> in{(const Tin*)a.in, a.in_tab, a.bs_in};
    const BatchPtr<Tout> out{(Tout*)a.out, a.out_tab, a.bs_out};
    hipLaunchKernelGGL(k, dim3((unsigned)blocks), dim3(C::NT), lds, a.stream, in, out, a.nx, (T)a.scale,
                       (const T*)a.tables, a.done, a.nplanes, a.batch);
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
    const int tile = a.tile ? a.tile : default_tile_y<T>(LOGN, a.nx);
    if (!tile_ok<T>(LOGN, tile)) return hipErrorInvalidValue;
    if (tile == 16) return yaxis_tile<Tin, Tout, T, LOGN, 16>(a);
    if constexpr (LOGN >= 5) {
        if constexpr (tile_lds_bytes(LOGN, 32, sizeof(T)) <= kLdsBytes)
            if (tile == 32) return yaxis_tile<Tin, Tout, T, LOGN, 32>(a);
        if constexpr (tile_lds_bytes(LOGN, 64, sizeof(T)) <= kLdsBytes)
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
