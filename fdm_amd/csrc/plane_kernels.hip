// Plane kernels (gfx950): transforms along x and y of whole z-planes held in
// LDS, and the refinement's forward transform with the FP64 residual fused
// into its load.
//
// Data layout in HBM: a field is dense [nz][ny][nx] (x fastest), interior
// points only, exactly the array LaplCube::solve receives (reference
// src/lapl_cube.h:80, src/tensor.h:192-205).  Batches of equal-shape problems
// are back to back (batch stride = nx*ny*nz elements) or addressed through a
// per-problem pointer table.  The transform is the DST-I on Dirichlet axes and
// the half-complex real DFT pair on periodic ones (template parameter KIND).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "dst_device.hpp"
#include "kernel_util.hpp"
#include "kernels.hpp"

namespace fdmgpu {

// ---------------------------------------------------------------------------
// plane kernel
// ---------------------------------------------------------------------------
template <int LOGN>
struct PlaneCfg {
    using S = DstShape<LOGN>;
    static constexpr int N = S::N;
    static constexpr int P = N + S::PAD;
    static constexpr int WORK = (N / 2) * S::N1;
    static constexpr int NT = WORK < 64 ? 64 : WORK;
    template <typename T>
    static constexpr size_t lds_bytes() {
        return sizeof(T) * (size_t)N * P;
    }
    template <typename T>
    static constexpr int waves() {
        return lds_waves_per_simd(lds_bytes<T>(), NT);
    }
};

// Global <-> LDS sweep over an n x n plane: thread tid owns column
// x = tid % N and rows r0, r0 + R, ... (r0 = tid / N, R = NT / N), so the
// global accesses of a wavefront are consecutive x and the row-dependent part
// of every address is a per-iteration constant (no per-element division).
template <int LOGN, int NT, int NN, int UNROLL = 64, typename F>
__device__ __forceinline__ void plane_sweep(const int tid, F&& f) {
    constexpr int N = 1 << LOGN, R = NT / N, ITER = (NN + R - 1) / R;
    static_assert(NT % N == 0, "a block covers whole rows");
    const int x = tid & (N - 1), r0 = tid >> LOGN;
    if (x < NN) {
#pragma unroll UNROLL
        for (int i = 0; i < ITER; ++i) {
            const int y = r0 + R * i;
            if (R * i + R <= NN || y < NN) f(i, y, x);
        }
    }
}

// The same sweep as a two-stage copy: every global load of the thread is
// issued into registers before the first store into LDS.  Left to itself the
// scheduler interleaves load, wait and store per row when the transform's
// register demand is high (FP64, N = 128), which puts one memory round trip
// per row on the critical path.
template <int LOGN, int NT, int NN, typename V, typename LD, typename ST>
__device__ __forceinline__ void plane_stage(const int tid, LD&& ld, ST&& st) {
    constexpr int N = 1 << LOGN, R = NT / N, ITER = (NN + R - 1) / R;
    static_assert(NT % N == 0, "a block covers whole rows");
    const int x = tid & (N - 1), r0 = tid >> LOGN;
    if (x < NN) {
        V v[ITER];
#pragma unroll
        for (int i = 0; i < ITER; ++i) {
            const int y = r0 + R * i;
            if (R * i + R <= NN || y < NN) v[i] = ld(i, y, x);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < ITER; ++i) {
            const int y = r0 + R * i;
            if (R * i + R <= NN || y < NN) st(i, y, x, v[i]);
        }
    }
}

// LDS slots of a sweep's elements: element i of thread tid is slot
// (r0 + R*i + o, x + o) of a PlaneLayout (x = tid % N, r0 = tid / N,
// R = NT / N).  The row swizzle f() has period N1 in the row, so rows a
// multiple of lcm(R, N1) apart share it: PER bases per thread, and every
// element is a base plus a compile-time offset (no per-element arithmetic).
template <int LOGN, int NT, int P, int N1, int O>
struct SweepSlots {
    static constexpr int N = 1 << LOGN, R = NT / N;
    static constexpr int PER = N1 / gcd_int(R, N1);
    using Lay = PlaneLayout<P, N1, true>;
    int b[PER];
    __device__ explicit SweepSlots(const int tid) {
        const int x = tid & (N - 1), r0 = tid >> LOGN;
#pragma unroll
        for (int j = 0; j < PER; ++j) b[j] = opaque(Lay::at(r0 + R * j + O, x + O));
    }
    __device__ __forceinline__ int operator()(const int i) const { return b[i % PER] + R * (i - i % PER) * P; }
};

// KIND (dst_device.hpp): kDst for Dirichlet planes of (N-1)^2 points,
// kRdftFwd / kRdftInv for periodic planes of N^2 points.
// ACC: 0 out = result; 1 out += result; 2 out = base + result with an FP32
// base (the refinement's first iterate), and out = base for problems flagged
// done (converged on that iterate).
template <typename Tin, typename Tout, typename T, int LOGN, int ACC, int KIND>
__global__ __launch_bounds__(PlaneCfg<LOGN>::NT)
__attribute__((amdgpu_waves_per_eu(PlaneCfg<LOGN>::template waves<T>()))) void plane_dst_kernel(
    const BatchPtr<const Tin> in, const BatchPtr<Tout> out, const T sx, const T sy, const T* __restrict__ tables,
    const int* __restrict__ done, const float* __restrict__ base, const long base_bs) {
    using C = PlaneCfg<LOGN>;
    using S = DstShape<LOGN>;
    using IO = PencilIO<KIND, LOGN>;
    using LX = PlaneLayout<C::P, S::N1, true>;
    using LY = PlaneLayout<C::P, S::N1, false>;
    constexpr int N = C::N, n = IO::n, o = IO::off, NT = C::NT;
    const int tid = threadIdx.x;
    const long plane_off = (long)blockIdx.x * n * n;
    const auto dst = as_global(out(blockIdx.y) + plane_off);
    [[maybe_unused]] const auto bp = as_global(ACC == 2 ? base + blockIdx.y * base_bs + plane_off : base);
    if (done && done[blockIdx.y]) {
        if constexpr (ACC == 2)
            plane_sweep<LOGN, NT, n>(tid, [&](int, int y, int x) { dst[y * n + x] = (Tout)bp[y * n + x]; });
        return;
    }
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* lds = reinterpret_cast<T*>(smem_raw);
    // the plane is read through a buffer resource: per element a lane offset
    // plus a compile-time one.  (The stores keep the natural global order and
    // read the permuted LDS slots: sweeping the LDS slots in order and
    // scattering rows on the global side issues fewer instructions but
    // measured 5-9% slower.)
    constexpr unsigned plane_pts = (unsigned)n * n;
    const BufRange<Tin> src(in(blockIdx.y) + plane_off, plane_pts * sizeof(Tin));
    const SweepSlots<LOGN, NT, C::P, S::N1, o> slot(tid);
    constexpr int R = NT / N;
    const int lane_in = (int)(((tid >> LOGN) * n + (tid & (N - 1))) * sizeof(Tin));

    DstLane<T, LOGN> L;
    L.init(tables, tid & (S::N1 - 1));
    if constexpr (KIND == kDst) {
        for (int i = tid; i < N; i += NT) {
            lds[LX::at(0, i)] = T(0);   // dummy row 0
            lds[LX::at(i, 0)] = T(0);   // dummy column 0
        }
    }
    plane_stage<LOGN, NT, n, Tin>(
        tid, [&](int i, int, int) { return src.load(lane_in, (int)(R * i * n * sizeof(Tin))); },
        [&](int i, int, int, Tin v) { lds[slot(i)] = (T)v; });
    __syncthreads();
    pencil_pairs<KIND, T, LOGN, N / 2, LX, NT>(lds, L, sx, tid);   // pencils = rows    (x)
    __syncthreads();
    // the y pass at N = 128 with per-lane slot tables (the x pass measured
    // 1-3% slower with them: its generic addressing pairs LDS accesses)
    if constexpr (KIND == kDst && LOGN == 7)
        dst_pairs128<T, C::P, N / 2, NT, true>(lds, L, sy, tid);   // pencils = columns (y)
    else
        pencil_pairs<KIND, T, LOGN, N / 2, LY, NT>(lds, L, sy, tid);
    __syncthreads();
    if constexpr (ACC == 2) {   // the FP32 base: every load in flight before the first store
        plane_stage<LOGN, NT, n, float>(
            tid, [&](int, int y, int x) { return bp[y * n + x]; },
            [&](int, int y, int x, float b) { dst[y * n + x] = (Tout)b + (Tout)lds[LX::at(IO::out(y), IO::out(x))]; });
    } else {
        plane_sweep<LOGN, NT, n>(tid, [&](int, int y, int x) {
            const T v = lds[LX::at(IO::out(y), IO::out(x))];
            if constexpr (ACC == 1) dst[y * n + x] += (Tout)v;
            else dst[y * n + x] = (Tout)v;
        });
    }
}

// ---------------------------------------------------------------------------
// Streaming plane kernel for FP64 planes of N = 128 (the plain LaplCube<double>
// solve at 127^3, periodic 128^3).  Such a plane takes 139 KB of LDS, so one
// block fills a CU; the one-plane-per-block kernel then leaves HBM idle while
// the block transforms and the CU idle while it loads.  Here a block walks
// planes q = blockIdx.x, blockIdx.x + gridDim.x, ... (about the same count per
// block: the grid is sized so), and while plane q is transformed the global
// loads of its next plane are in flight in registers: the first kStreamSplit
// rows of a thread's share issued before the x pass, the rest before the y
// pass.  After the y pass the outputs are read into registers and, once every
// wave has read its share, stored kStreamChunk rows at a time between the
// next plane's staging writes.  (A wave issues its memory instructions in
// order and stalls while the memory pipeline is full: a whole plane of loads
// or of stores issued at once held every wave for about 4k cycles per plane,
// measured with s_memtime stamps.)  The plane sits in LDS in the
// b64-conflict-free PlaneLayoutB64 (pitch kStreamPitch, jobs dealt to pencils
// out of order).  Each pencil goes through plane_dst_kernel's transform;
// results agree with that kernel's to FP64 rounding (2e-14 relative at
// 127^3), not bitwise.
// ---------------------------------------------------------------------------
constexpr int kStreamPitch = 130;
constexpr int kStreamSplit = 16;   // rows (of 32 per thread) prefetched before the x pass
constexpr int kStreamChunk = 8;    // rows stored between staging writes

template <int KIND>
__global__ __launch_bounds__(PlaneCfg<7>::NT) __attribute__((amdgpu_waves_per_eu(2))) void plane_stream_kernel(
    const BatchPtr<const double> in, const BatchPtr<double> out, const double sx, const double sy,
    const double* __restrict__ tables, const int nplanes, const int total) {
    constexpr int LOGN = 7;
    using C = PlaneCfg<LOGN>;
    using S = DstShape<LOGN>;
    using IO = PencilIO<KIND, LOGN>;
    constexpr int P = kStreamPitch;   // conflict-free b64 layout (PlaneLayoutB64)
    using LX = PlaneLayoutB64<P, true>;
    using LY = PlaneLayoutB64<P, false>;
    constexpr int N = C::N, n = IO::n, o = IO::off, NT = C::NT;
    constexpr int R = NT / N, ITER = (n + R - 1) / R;
    static_assert(NT % N == 0, "a block covers whole rows");
    const int tid = threadIdx.x;
    int q = blockIdx.x;
    if (q >= total) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double* lds = reinterpret_cast<double*>(smem_raw);
    constexpr unsigned plane_pts = (unsigned)n * n;
    // Thread (r0, x0) loads, stages and stores column x0 of rows ITER r0 + i,
    // i < ITER: a row group r0 is two whole waves, so r0 is wave-uniform and
    // every row-dependent offset is a compile-time constant within a group.
    // Lanes outside the plane (x0 >= n, or the rows >= n of the last group)
    // load in bounds or past the buffer's range (which reads 0), stage into
    // the pad column, and store past the range (which is dropped).
    const int x0 = tid & (N - 1), r0 = __builtin_amdgcn_readfirstlane(tid >> LOGN);
    const int slot0 = opaque(LX::at(ITER * r0 + o, x0 + o));   // element i: slot0 + i P
    const int lane_in = (int)((ITER * r0 * n + x0) * sizeof(double));

    // the DST's shifted sines come from an LDS copy behind the plane (the
    // prefetch takes their registers)
    using Lane = DstLaneLdsSine<double, LOGN>;
    double* sines = lds + N * P;
    Lane L;
    Lane::fill_sines(tables, sines, tid, NT);
    L.init(tables, sines, tid & (S::N1 - 1));
    if constexpr (KIND == kDst) {   // the dummy row and column stay 0 across planes
        for (int i = tid; i < N; i += NT) {
            lds[LX::at(0, i)] = 0.0;
            lds[LX::at(i, 0)] = 0.0;
        }
    }
    double v[ITER];
    // (problem, plane) of a plane index, wave-uniform: in SGPRs, so the
    // buffer resource and the store base need no per-lane handling
    auto split = [&](const int qq, int& p, int& z) {
        p = __builtin_amdgcn_readfirstlane(qq / nplanes);
        z = __builtin_amdgcn_readfirstlane(qq - p * nplanes);
    };
    // rows [LO, HI) of the thread's share of plane qq into v
    auto load_rows = [&](const int qq, auto loc, auto hic) {
        constexpr int LO = decltype(loc)::value, HI = decltype(hic)::value;
        int p, z;
        split(qq, p, z);
        const BufRange<double> src(wave_uniform(in.uniform(p)) + (long)z * plane_pts, plane_pts * sizeof(double));
#pragma unroll
        for (int i = LO; i < HI; ++i) v[i] = src.load(lane_in, (int)(i * n * sizeof(double)));
    };
    using R0 = std::integral_constant<int, 0>;
    using RH = std::integral_constant<int, kStreamSplit>;
    using R1 = std::integral_constant<int, ITER>;
    auto load_plane = [&](const int qq) { load_rows(qq, R0{}, R1{}); };
    auto stage_rows = [&](auto loc, auto hic) {
        constexpr int LO = decltype(loc)::value, HI = decltype(hic)::value;
#pragma unroll
        for (int i = LO; i < HI; ++i)   // a row >= n would land on the sines behind the plane
            if (ITER * (R - 1) + i < n || ITER * r0 + i < n) lds[slot0 + i * P] = v[i];
    };
    auto stage = [&] { stage_rows(std::integral_constant<int, 0>{}, std::integral_constant<int, ITER>{}); };
    // Staging sits at the end of the loop body, behind the store phase, and
    // the first plane is staged before the loop, after which every load so
    // far has completed: nothing the loop reads from global memory is then
    // pending at the loop head.  (vmcnt retires in order; a value pending
    // there, such as the lane constants loaded above, gets its wait inside
    // the loop, where it also waits for the whole prefetch in flight.)
    // Within an iteration the prefetch is issued before the store phase's
    // stores, so staging waits for the prefetch only, never for those stores.
    load_plane(q);
    stage();
    wait_vmem_all();
    for (;;) {
        __syncthreads();
        const int qn = q + (int)gridDim.x;
        // the next plane's first rows, in flight until the next staging; the
        // rest are issued after the x pass (issuing a whole plane at once
        // stalls the wave while the memory pipeline takes them)
        if (qn < total) load_rows(qn, R0{}, RH{});
        __builtin_amdgcn_sched_barrier(0);
        // the passes' addresses are recomputed per plane from a thread index
        // the compiler cannot hoist out of the loop: kept live across the
        // loop they would take the registers the prefetch needs
        int t1 = tid;
        asm volatile("" : "+v"(t1));
        pencil_pairs<KIND, double, LOGN, N / 2, LX, NT, Lane>(lds, L, sx, t1);   // pencils = rows    (x)
        __syncthreads();
        if constexpr (kStreamSplit < ITER)
            if (qn < total) load_rows(qn, RH{}, R1{});
        __builtin_amdgcn_sched_barrier(0);
        int t2 = tid;
        asm volatile("" : "+v"(t2));
        if constexpr (KIND == kDst)
            dst_pairs128<double, P, N / 2, NT, true, Lane, true>(lds, L, sy, t2);   // pencils = columns (y)
        else
            pencil_pairs<KIND, double, LOGN, N / 2, LY, NT, Lane>(lds, L, sy, t2);
        __syncthreads();
        // store phase: the plane's outputs into registers, then, once every
        // wave has read its share, their global stores interleaved with the
        // next plane's staging (a plane of stores issued at once stalls the
        // wave while the memory pipeline takes them; staged in between, the
        // LDS writes proceed meanwhile)
        int t3 = tid;
        asm volatile("" : "+v"(t3));
        int p, z;
        split(q, p, z);
        const BufRange<double> dst(wave_uniform(out.uniform(p)) + (long)z * plane_pts, plane_pts * sizeof(double));
        // Every lane stores every row, with no branch: a lane outside the
        // plane (x or y >= n) reads a slot inside it and stores beyond the
        // buffer's range, which the hardware drops.  (A branch around the
        // stores would make the compiler assume at staging that they might
        // not have been issued, and count the waits for the prefetch without
        // them: staging would then wait for the stores.)
        const int xo = t3 & (N - 1);
        const int xs = opaque(IO::out(xo < n ? xo : n - 1));
        const int lane_out = xo < n ? (int)((ITER * r0 * n + xo) * sizeof(double)) : (int)(plane_pts * sizeof(double));
        double o[ITER];
        // the row part of each LDS slot is wave-uniform (scalar arithmetic);
        // a lane adds its column slot
#pragma unroll
        for (int i = 0; i < ITER; ++i) {
            const int y = ITER * r0 + i, ys = y < n - 1 ? y : n - 1;
            o[i] = lds[LX::at(IO::out(ys), 0) + xs];
        }
        auto store_rows = [&](auto loc, auto hic) {
            constexpr int LO = decltype(loc)::value, HI = decltype(hic)::value;
#pragma unroll
            for (int i = LO; i < HI; ++i) dst.store(o[i], lane_out, (int)(i * n * sizeof(double)));
        };
        if (qn >= total) {
            store_rows(std::integral_constant<int, 0>{}, std::integral_constant<int, ITER>{});
            break;
        }
        q = qn;
        __syncthreads();   // every wave has read the plane before the next one is staged
        static_assert(ITER % kStreamChunk == 0, "whole chunks");
        static_for<0, ITER / kStreamChunk>([&](auto cc) {
            constexpr int c = decltype(cc)::value;
            using Lo = std::integral_constant<int, c * kStreamChunk>;
            using Hi = std::integral_constant<int, (c + 1) * kStreamChunk>;
            store_rows(Lo{}, Hi{});
            stage_rows(Lo{}, Hi{});
        });
    }
}

// ---------------------------------------------------------------------------
// Streaming plane kernel for FP32 planes of N = 128 (LaplCube<float> and the
// refinement's FP32 inner solves at 127^3; periodic 128^3): the scheme of
// plane_stream_kernel for a plane of floats.  Such a plane takes 70 KB of LDS,
// so two blocks share a CU (the grid is two per CU) and a block holds at most
// 128 registers per lane: the lane constants' shifted sines are read from LDS
// and a thread's 32 prefetched rows take 32 of them.  The plane sits in the
// swizzled PlaneLayout of plane_dst_kernel; a thread's rows are ITER
// consecutive ones starting on a multiple of the swizzle period, so each
// row's slot is the thread's first plus a compile-time constant.  Problems
// flagged in `done` are skipped (their planes are not read or written).  The
// y pass is the generic column transform (plane_dst_kernel's slot-table
// variant of it would spill here), so results agree with plane_dst_kernel's
// to float rounding, not bitwise.
// ---------------------------------------------------------------------------
constexpr int kStream32Split = 16;
constexpr int kStream32Chunk = 8;

template <int KIND>
__global__ __launch_bounds__(PlaneCfg<7>::NT) __attribute__((amdgpu_waves_per_eu(4))) void plane_stream32_kernel(
    const BatchPtr<const float> in, const BatchPtr<float> out, const float sx, const float sy,
    const float* __restrict__ tables, const int* __restrict__ done, const int nplanes, const int total) {
    constexpr int LOGN = 7;
    using C = PlaneCfg<LOGN>;
    using S = DstShape<LOGN>;
    using IO = PencilIO<KIND, LOGN>;
    using LX = PlaneLayout<C::P, S::N1, true>;
    using LY = PlaneLayout<C::P, S::N1, false>;
    using Sw = typename LX::Sw;
    constexpr int N = C::N, P = C::P, n = IO::n, o = IO::off, NT = C::NT;
    constexpr int R = NT / N, ITER = (n + R - 1) / R;
    static_assert(NT % N == 0 && ITER % S::N1 == 0, "row groups start on a swizzle period");
    const int tid = threadIdx.x;
    const int G = (int)gridDim.x;
    // a plane of a problem flagged done is skipped (wave-uniform: scalar load)
    auto skipped = [&](const int qq) {
        if (!done) return false;
        using CI = const __attribute__((address_space(4))) int*;
        return ((CI)done)[__builtin_amdgcn_readfirstlane(qq / nplanes)] != 0;
    };
    auto next = [&](int qq) {
        do qq += G;
        while (qq < total && skipped(qq));
        return qq;
    };
    int q = blockIdx.x;
    if (q < total && skipped(q)) q = next(q);
    if (q >= total) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float* lds = reinterpret_cast<float*>(smem_raw);
    constexpr unsigned plane_pts = (unsigned)n * n;
    // thread (r0, x0): column x0 of rows ITER r0 + i, i < ITER (r0 wave-uniform)
    const int x0 = tid & (N - 1), r0 = __builtin_amdgcn_readfirstlane(tid >> LOGN);
    const int slot0 = opaque(LX::at(ITER * r0 + o, x0 + o));
    const int lane_in = (int)((ITER * r0 * n + x0) * sizeof(float));
    auto slot = [&](const int i) { return slot0 + i * P + Sw::f(i + o) - Sw::f(o); };

    using Lane = DstLaneLdsSine<float, LOGN>;
    float* sines = lds + N * P;
    Lane L;
    Lane::fill_sines(tables, sines, tid, NT);
    L.init(tables, sines, tid & (S::N1 - 1));
    if constexpr (KIND == kDst) {   // the dummy row and column stay 0 across planes
        for (int i = tid; i < N; i += NT) {
            lds[LX::at(0, i)] = 0.f;
            lds[LX::at(i, 0)] = 0.f;
        }
    }
    float v[ITER];
    auto split = [&](const int qq, int& p, int& z) {
        p = __builtin_amdgcn_readfirstlane(qq / nplanes);
        z = __builtin_amdgcn_readfirstlane(qq - p * nplanes);
    };
    auto load_rows = [&](const int qq, auto loc, auto hic) {
        constexpr int LO = decltype(loc)::value, HI = decltype(hic)::value;
        int p, z;
        split(qq, p, z);
        const BufRange<float> src(wave_uniform(in.uniform(p)) + (long)z * plane_pts, plane_pts * sizeof(float));
#pragma unroll
        for (int i = LO; i < HI; ++i) v[i] = src.load(lane_in, (int)(i * n * sizeof(float)));
    };
    using R0 = std::integral_constant<int, 0>;
    using RH = std::integral_constant<int, kStream32Split>;
    using R1 = std::integral_constant<int, ITER>;
    auto stage_rows = [&](auto loc, auto hic) {
        constexpr int LO = decltype(loc)::value, HI = decltype(hic)::value;
#pragma unroll
        for (int i = LO; i < HI; ++i)   // a row >= n would land on the sines behind the plane
            if (ITER * (R - 1) + i < n || ITER * r0 + i < n) lds[slot(i)] = v[i];
    };
    // the loop's order of loads, waits and stores is plane_stream_kernel's
    load_rows(q, R0{}, R1{});
    stage_rows(R0{}, R1{});
    wait_vmem_all();
    for (;;) {
        __syncthreads();
        const int qn = next(q);
        if (qn < total) load_rows(qn, R0{}, RH{});
        __builtin_amdgcn_sched_barrier(0);
        int t1 = tid;
        asm volatile("" : "+v"(t1));
        pencil_pairs<KIND, float, LOGN, N / 2, LX, NT, Lane>(lds, L, sx, t1);   // pencils = rows    (x)
        __syncthreads();
        if constexpr (kStream32Split < ITER)
            if (qn < total) load_rows(qn, RH{}, R1{});
        __builtin_amdgcn_sched_barrier(0);
        int t2 = tid;
        asm volatile("" : "+v"(t2));
        // (dst_pairs128's slot tables would take registers the prefetch holds)
        pencil_pairs<KIND, float, LOGN, N / 2, LY, NT, Lane>(lds, L, sy, t2);   // pencils = columns (y)
        __syncthreads();
        int t3 = tid;
        asm volatile("" : "+v"(t3));
        int p, z;
        split(q, p, z);
        const BufRange<float> dst(wave_uniform(out.uniform(p)) + (long)z * plane_pts, plane_pts * sizeof(float));
        // every lane stores every row, with no branch: a lane outside the
        // plane reads a slot inside it and stores beyond the buffer's range
        const int xo = t3 & (N - 1);
        const int xs = opaque(IO::out(xo < n ? xo : n - 1));
        const int lane_out = xo < n ? (int)((ITER * r0 * n + xo) * sizeof(float)) : (int)(plane_pts * sizeof(float));
        float ov[ITER];
#pragma unroll
        for (int i = 0; i < ITER; ++i) {
            const int y = ITER * r0 + i, ys = y < n - 1 ? y : n - 1;
            ov[i] = lds[LX::at(IO::out(ys), 0) + xs];
        }
        auto store_rows = [&](auto loc, auto hic) {
            constexpr int LO = decltype(loc)::value, HI = decltype(hic)::value;
#pragma unroll
            for (int i = LO; i < HI; ++i) dst.store(ov[i], lane_out, (int)(i * n * sizeof(float)));
        };
        if (qn >= total) {
            store_rows(R0{}, R1{});
            break;
        }
        q = qn;
        __syncthreads();   // every wave has read the plane before the next one is staged
        static_assert(ITER % kStream32Chunk == 0, "whole chunks");
        static_for<0, ITER / kStream32Chunk>([&](auto cc) {
            constexpr int c = decltype(cc)::value;
            using Lo = std::integral_constant<int, c * kStream32Chunk>;
            using Hi = std::integral_constant<int, (c + 1) * kStream32Chunk>;
            store_rows(Lo{}, Hi{});
            stage_rows(Lo{}, Hi{});
        });
    }
}

// ---------------------------------------------------------------------------
// Refinement: forward DST of the FP64 residual r = f - A u, fused into the
// plane load.  A is the 7-point operator the solve inverts (1/h^2 per axis:
// ix, iy, iz).  Each block also writes (sum r^2, sum f^2) of its plane to
// partial[p][z] for the convergence check.  Planes are dealt to XCDs in
// contiguous z-slabs (workgroups go round-robin over the 8 XCDs), so a
// block's z-1 / z+1 neighbour planes are read by the same XCD close in time
// and come from its L2.
// ---------------------------------------------------------------------------
constexpr int kXcds = 8;

__device__ __forceinline__ int xcd_slab_plane(int bid, int nplanes) {
    const int per = (nplanes + kXcds - 1) / kXcds;
    return (bid % kXcds) * per + bid / kXcds;
}

template <typename Tu>
__device__ __forceinline__ double residual_at(const global_ptr<const double> f, const global_ptr<const Tu> u, const long c,
                                              const int x, const int y, const int z, const int nx, const int ny,
                                              const int nz, const double ix, const double iy, const double iz,
                                              double& fv) {
    const long pl = (long)nx * ny;
    const double uc = (double)u[c];
    const double xm = x > 0 ? (double)u[c - 1] : 0.0, xp = x < nx - 1 ? (double)u[c + 1] : 0.0;
    const double ym = y > 0 ? (double)u[c - nx] : 0.0, yp = y < ny - 1 ? (double)u[c + nx] : 0.0;
    const double zm = z > 0 ? (double)u[c - pl] : 0.0, zp = z < nz - 1 ? (double)u[c + pl] : 0.0;
    fv = f[c];
    return fv - ((xm - 2.0 * uc + xp) * ix + (ym - 2.0 * uc + yp) * iy + (zm - 2.0 * uc + zp) * iz);
}

template <int NT>
__device__ __forceinline__ void block_sum2_to(double rr, double ff, double* __restrict__ out,
                                              double* __restrict__ red) {
    // wave reduction, then the waves' partials in fixed order
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        rr += __shfl_xor(rr, d, 64);
        ff += __shfl_xor(ff, d, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[2 * w] = rr;
        red[2 * w + 1] = ff;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int i = 0; i < (NT + 63) / 64; ++i) {
            a += red[2 * i];
            b += red[2 * i + 1];
        }
        out[0] = a;
        out[1] = b;
    }
}

template <typename Tu, int LOGN>
__global__ __launch_bounds__(PlaneCfg<LOGN>::NT)
__attribute__((amdgpu_waves_per_eu(PlaneCfg<LOGN>::template waves<float>()))) void plane_resid_dst_kernel(
    const BatchPtr<const double> f, const BatchPtr<const Tu> u, float* __restrict__ out,
    double* __restrict__ partial, const long bs, const int nplanes, const double ix, const double iy,
    const double iz, const float sx, const float sy, const float* __restrict__ tables,
    const int* __restrict__ done) {
    using C = PlaneCfg<LOGN>;
    using S = DstShape<LOGN>;
    using LX = PlaneLayout<C::P, S::N1, true>;
    using LY = PlaneLayout<C::P, S::N1, false>;
    constexpr int N = C::N, n = N - 1, NT = C::NT;
    const int p = blockIdx.y;
    const int z = xcd_slab_plane(blockIdx.x, nplanes);
    if (z >= nplanes || (done && done[p])) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float* lds = reinterpret_cast<float*>(smem_raw);
    double* red = reinterpret_cast<double*>(lds + N * C::P + 4);   // 2 doubles per wave
    const int tid = threadIdx.x;

    DstLane<float, LOGN> L;
    L.init(tables, tid & (S::N1 - 1));
    for (int i = tid; i < N; i += NT) {
        lds[LX::at(0, i)] = 0.f;
        lds[LX::at(i, 0)] = 0.f;
    }
    const auto fp = as_global(f(p));
    const auto up = as_global(u(p));
    double rr = 0.0, ff = 0.0;
    plane_sweep<LOGN, NT, n, 2>(tid, [&](int, int y, int x) {   // 8 loads per point: unroll little (registers)
        double fv;
        const double r =
            residual_at<Tu>(fp, up, (long)z * n * n + y * n + x, x, y, z, n, n, nplanes, ix, iy, iz, fv);
        rr += r * r;
        ff += fv * fv;
        lds[LX::at(y + 1, x + 1)] = (float)r;
    });
    block_sum2_to<NT>(rr, ff, partial + ((long)p * nplanes + z) * 2, red);
    __syncthreads();
    dst_pairs<float, LOGN, N / 2, LX, NT>(lds, L, sx, tid);
    __syncthreads();
    dst_pairs<float, LOGN, N / 2, LY, NT>(lds, L, sy, tid);
    __syncthreads();
    float* dst = out + p * bs + (long)z * n * n;
    plane_sweep<LOGN, NT, n>(tid, [&](int, int y, int x) { dst[y * n + x] = lds[LX::at(S::sigma(y + 1), S::sigma(x + 1))]; });
}

template <typename Tin, typename Tout, typename T, int LOGN, int ACC, int KIND>
static hipError_t plane_k(const PlaneArgs& a) {
    using C = PlaneCfg<LOGN>;
    auto k = plane_dst_kernel<Tin, Tout, T, LOGN, ACC, KIND>;
    const size_t lds = C::template lds_bytes<T>();
    hipError_t e = allow_lds(k, lds);
    if (e != hipSuccess) return e;
    const BatchPtr<const Tin> in{(const Tin*)a.in, a.in_tab, a.bs_in};
    const BatchPtr<Tout> out{(Tout*)a.out, a.out_tab, a.bs_out};
    hipLaunchKernelGGL(k, dim3(a.nplanes, a.batch), dim3(C::NT), lds, a.stream, in, out, (T)a.sx, (T)a.sy,
                       (const T*)a.tables, a.done, a.base, a.bs_out);
    return hipGetLastError();
}

// Compute units of the current device (cached per device).
static int device_cus() {
    static int cus[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (!cus[dev]) {
        int c = 0;
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c < 1) c = 256;
        cus[dev] = c;
    }
    return cus[dev];
}

template <int KIND>
static hipError_t plane_stream_k(const PlaneArgs& a) {
    using C = PlaneCfg<7>;
    auto k = plane_stream_kernel<KIND>;
    const size_t lds = sizeof(double) * (size_t)C::N * (kStreamPitch + 1);   // plane + sine table
    hipError_t e = allow_lds(k, lds);
    if (e != hipSuccess) return e;
    const long total = (long)a.nplanes * a.batch;
    if (total < 1 || total > 0x7fffffffL) return hipErrorInvalidValue;
    // one block per CU; every block gets ceil(total / CUs) planes or one fewer
    const long per = (total + device_cus() - 1) / device_cus();
    const int grid = (int)((total + per - 1) / per);
    const BatchPtr<const double> in{(const double*)a.in, a.in_tab, a.bs_in};
    const BatchPtr<double> out{(double*)a.out, a.out_tab, a.bs_out};
    hipLaunchKernelGGL(k, dim3(grid), dim3(C::NT), lds, a.stream, in, out, a.sx, a.sy, (const double*)a.tables,
                       a.nplanes, (int)total);
    return hipGetLastError();
}

template <int KIND>
static hipError_t plane_stream32_k(const PlaneArgs& a) {
    using C = PlaneCfg<7>;
    auto k = plane_stream32_kernel<KIND>;
    const size_t lds = sizeof(float) * (size_t)C::N * (C::P + 1);   // plane + sine table
    hipError_t e = allow_lds(k, lds);
    if (e != hipSuccess) return e;
    const long total = (long)a.nplanes * a.batch;
    if (total < 1 || total > 0x7fffffffL) return hipErrorInvalidValue;
    // two blocks per CU; every block gets ceil(total / blocks) planes or one fewer
    const long blocks = 2L * device_cus();
    const long per = (total + blocks - 1) / blocks;
    const int grid = (int)((total + per - 1) / per);
    const BatchPtr<const float> in{(const float*)a.in, a.in_tab, a.bs_in};
    const BatchPtr<float> out{(float*)a.out, a.out_tab, a.bs_out};
    hipLaunchKernelGGL(k, dim3(grid), dim3(C::NT), lds, a.stream, in, out, (float)a.sx, (float)a.sy,
                       (const float*)a.tables, a.done, a.nplanes, (int)total);
    return hipGetLastError();
}

template <typename Tin, typename Tout, typename T, int LOGN>
static hipError_t plane_t(const PlaneArgs& a) {
    if (a.base && !a.accum) return hipErrorInvalidValue;
    if (a.stream_planes) {   // planes of N = 128: FP64 without done flags, FP32 in and out
        if constexpr (LOGN == 7 && std::is_same_v<Tin, double> && std::is_same_v<Tout, double> &&
                      std::is_same_v<T, double>) {
            if (a.accum || a.base || a.done) return hipErrorInvalidValue;
            if (a.kind == kDst) return plane_stream_k<kDst>(a);
            if (a.kind == kRdftFwd) return plane_stream_k<kRdftFwd>(a);
            if (a.kind == kRdftInv) return plane_stream_k<kRdftInv>(a);
            return hipErrorInvalidValue;
        }
        if constexpr (LOGN == 7 && std::is_same_v<Tin, float> && std::is_same_v<Tout, float> &&
                      std::is_same_v<T, float>) {
            if (!a.accum) {
                if (a.kind == kDst) return plane_stream32_k<kDst>(a);
                if (a.kind == kRdftFwd) return plane_stream32_k<kRdftFwd>(a);
                if (a.kind == kRdftInv) return plane_stream32_k<kRdftInv>(a);
                return hipErrorInvalidValue;
            }
        }
        // the refinement's other FP32 planes (FP64 in, or FP64 out with an
        // FP32 base) take plane_dst_kernel below
        if constexpr (!std::is_same_v<T, float>) return hipErrorInvalidValue;
    }
    if (a.kind == kDst) {
        if constexpr (std::is_same_v<Tout, double> && !std::is_same_v<Tin, double>) {
            if (a.base) return plane_k<Tin, Tout, T, LOGN, 2, kDst>(a);
        }
        if (a.base) return hipErrorInvalidValue;
        return a.accum ? plane_k<Tin, Tout, T, LOGN, 1, kDst>(a) : plane_k<Tin, Tout, T, LOGN, 0, kDst>(a);
    }
    if constexpr (std::is_same_v<Tin, T> && std::is_same_v<Tout, T>) {
        if (a.accum) return hipErrorInvalidValue;
        if (a.kind == kRdftFwd) return plane_k<Tin, Tout, T, LOGN, 0, kRdftFwd>(a);
        if (a.kind == kRdftInv) return plane_k<Tin, Tout, T, LOGN, 0, kRdftInv>(a);
    }
    return hipErrorInvalidValue;
}

template <typename Tu, int LOGN>
static hipError_t plane_resid_t(const PlaneResidArgs& a) {
    using C = PlaneCfg<LOGN>;
    auto k = plane_resid_dst_kernel<Tu, LOGN>;
    const size_t lds = C::template lds_bytes<float>() + 16 + 2 * sizeof(double) * ((C::NT + 63) / 64);
    hipError_t e = allow_lds(k, lds);
    if (e != hipSuccess) return e;
    const int per = (a.nplanes + kXcds - 1) / kXcds;
    const BatchPtr<const double> f{a.f, a.f_tab, a.bs};
    const BatchPtr<const Tu> u{(const Tu*)a.u, a.u_tab, a.bs};
    hipLaunchKernelGGL(k, dim3(per * kXcds, a.batch), dim3(C::NT), lds, a.stream, f, u, a.out, a.partial, a.bs,
                       a.nplanes, a.ix, a.iy, a.iz, (float)a.sx, (float)a.sy, (const float*)a.tables, a.done);
    return hipGetLastError();
}

template <typename Tu>
hipError_t launch_plane_resid(int logn, const PlaneResidArgs& a) {
    switch (logn) {
        case 1: return plane_resid_t<Tu, 1>(a);
        case 2: return plane_resid_t<Tu, 2>(a);
        case 3: return plane_resid_t<Tu, 3>(a);
        case 4: return plane_resid_t<Tu, 4>(a);
        case 5: return plane_resid_t<Tu, 5>(a);
        case 6: return plane_resid_t<Tu, 6>(a);
        case 7: return plane_resid_t<Tu, 7>(a);
        default: return hipErrorInvalidValue;
    }
}
template hipError_t launch_plane_resid<double>(int, const PlaneResidArgs&);

template <typename Tin, typename Tout, typename T>
hipError_t launch_plane(int logn, const PlaneArgs& a) {
    switch (logn) {
        case 1: return plane_t<Tin, Tout, T, 1>(a);
        case 2: return plane_t<Tin, Tout, T, 2>(a);
        case 3: return plane_t<Tin, Tout, T, 3>(a);
        case 4: return plane_t<Tin, Tout, T, 4>(a);
        case 5: return plane_t<Tin, Tout, T, 5>(a);
        case 6: return plane_t<Tin, Tout, T, 6>(a);
        case 7: return plane_t<Tin, Tout, T, 7>(a);
        default: return hipErrorInvalidValue;
    }
}

template <typename T>
bool plane_fits(int logn) {
    if (logn < 1 || logn > 7) return false;
    size_t N = (size_t)1 << logn, N1 = (size_t)1 << (logn / 2);
    return sizeof(T) * (N * (N + N1)) <= 160 * 1024;
}

bool plane_stream_ok(int logn, bool) { return logn == 7; }

template hipError_t launch_plane<float, float, float>(int, const PlaneArgs&);
template hipError_t launch_plane<double, double, double>(int, const PlaneArgs&);
template hipError_t launch_plane<double, float, float>(int, const PlaneArgs&);
template hipError_t launch_plane<float, double, float>(int, const PlaneArgs&);
template bool plane_fits<float>(int);
template bool plane_fits<double>(int);

}  // namespace fdmgpu
