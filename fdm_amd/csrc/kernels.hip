// HIP kernels for the fast Poisson solve (gfx950) and their host launchers.
//
// Data layout in HBM: a field is dense [nz][ny][nx] (x fastest), interior
// points only, exactly the array LaplCube::solve receives (reference
// src/lapl_cube.h:80, src/tensor.h:192-205).  Batches of equal-shape problems
// are laid out back to back (batch stride = nx*ny*nz elements).
//
// Kernels (the transform is the DST-I on Dirichlet axes and the half-complex
// real DFT pair on periodic ones, template parameter KIND):
//   plane_dst_kernel   transforms along x and y of whole z-planes held in LDS
//                      (square planes, N <= 128): one HBM read + one write.
//   xaxis_dst_kernel   transform along x of a tile of TILE rows (any N <= 512).
//   yaxis_dst_kernel   transform along y of a tile of TILE columns of one plane.
//   zcr_kernel         constant-coefficient cyclic reduction along z for every
//                      (kx,ky) mode (the tridiagonal replacing the reference's
//                      z transform pair + divide).
//   zcrp_kernel        the same for a periodic z axis (circulant systems).
//   rowcr_kernel       general-n cyclic reduction along x with per-row
//                      coefficients (LaplRect's gtsv step).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "dst_device.hpp"
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
};

// KIND (dst_device.hpp): kDst for Dirichlet planes of (N-1)^2 points,
// kRdftFwd / kRdftInv for periodic planes of N^2 points.
template <typename Tin, typename Tout, typename T, int LOGN, bool ACCUM, int KIND>
__global__ __launch_bounds__(PlaneCfg<LOGN>::NT) void plane_dst_kernel(
    const Tin* __restrict__ in, Tout* __restrict__ out, const long bs_in, const long bs_out, const T sx,
    const T sy, const T* __restrict__ tables, const int* __restrict__ done) {
    using C = PlaneCfg<LOGN>;
    using S = DstShape<LOGN>;
    using IO = PencilIO<KIND, LOGN>;
    using LX = PlaneLayout<C::P, S::N1, true>;
    using LY = PlaneLayout<C::P, S::N1, false>;
    if (done && done[blockIdx.y]) return;
    constexpr int N = C::N, n = IO::n, o = IO::off, NT = C::NT;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* lds = reinterpret_cast<T*>(smem_raw);
    const int tid = threadIdx.x;
    const long plane_off = (long)blockIdx.x * n * n;
    const Tin* src = in + blockIdx.y * bs_in + plane_off;
    Tout* dst = out + blockIdx.y * bs_out + plane_off;

    DstLane<T, LOGN> L;
    L.init(tables, tid & (S::N1 - 1));
    if constexpr (KIND == kDst) {
        for (int i = tid; i < N; i += NT) {
            lds[LX::at(0, i)] = T(0);   // dummy row 0
            lds[LX::at(i, 0)] = T(0);   // dummy column 0
        }
    }
    for (int idx = tid; idx < n * n; idx += NT) {
        const int y = idx / n, x = idx - y * n;
        lds[LX::at(y + o, x + o)] = (T)src[idx];
    }
    __syncthreads();
    pencil_pairs<KIND, T, LOGN, N / 2, LX, NT>(lds, L, sx, tid);   // pencils = rows    (x)
    __syncthreads();
    pencil_pairs<KIND, T, LOGN, N / 2, LY, NT>(lds, L, sy, tid);   // pencils = columns (y)
    __syncthreads();
    for (int idx = tid; idx < n * n; idx += NT) {
        const int y = idx / n, x = idx - y * n;
        const T v = lds[LX::at(IO::out(y), IO::out(x))];
        if constexpr (ACCUM) dst[idx] += (Tout)v;
        else dst[idx] = (Tout)v;
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
__device__ __forceinline__ double residual_at(const double* __restrict__ f, const Tu* __restrict__ u, const long c,
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
__global__ __launch_bounds__(PlaneCfg<LOGN>::NT) void plane_resid_dst_kernel(
    const double* __restrict__ f, const Tu* __restrict__ u, float* __restrict__ out, double* __restrict__ partial,
    const long bs, const int nplanes, const double ix, const double iy, const double iz, const float sx,
    const float sy, const float* __restrict__ tables, const int* __restrict__ done) {
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
    double rr = 0.0, ff = 0.0;
    for (int idx = tid; idx < n * n; idx += NT) {
        const int y = idx / n, x = idx - y * n;
        double fv;
        const double r = residual_at<Tu>(f + p * bs, u + p * bs, (long)z * n * n + idx, x, y, z, n, n, nplanes, ix,
                                         iy, iz, fv);
        rr += r * r;
        ff += fv * fv;
        lds[LX::at(y + 1, x + 1)] = (float)r;
    }
    block_sum2_to<NT>(rr, ff, partial + ((long)p * nplanes + z) * 2, red);
    __syncthreads();
    dst_pairs<float, LOGN, N / 2, LX, NT>(lds, L, sx, tid);
    __syncthreads();
    dst_pairs<float, LOGN, N / 2, LY, NT>(lds, L, sy, tid);
    __syncthreads();
    float* dst = out + p * bs + (long)z * n * n;
    for (int idx = tid; idx < n * n; idx += NT) {
        const int y = idx / n, x = idx - y * n;
        dst[idx] = lds[LX::at(S::sigma(y + 1), S::sigma(x + 1))];
    }
}

// ---------------------------------------------------------------------------
// axis-tile kernels (planes that do not fit LDS, or non-square planes)
// ---------------------------------------------------------------------------
template <int LOGN, typename T>
struct TileCfg {
    using S = DstShape<LOGN>;
    static constexpr int N = S::N;
    // pencils per block: 32, or 16 when a 32-wide y tile would not fit LDS
    static constexpr bool BIG = (size_t)N * (32 + S::N1) * sizeof(T) > 160 * 1024;
    static constexpr int TILE = BIG ? 16 : 32;
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
template <typename Tin, typename Tout, typename T, int LOGN, bool ACCUM, int KIND>
__global__ __launch_bounds__((TileCfg<LOGN, T>::NT)) void xaxis_dst_kernel(
    const Tin* __restrict__ in, Tout* __restrict__ out, const long nrows, const long bs_in, const long bs_out,
    const T scale, const T* __restrict__ tables, const int* __restrict__ done) {
    using C = TileCfg<LOGN, T>;
    using S = DstShape<LOGN>;
    using IO = PencilIO<KIND, LOGN>;
    using LX = PlaneLayout<C::P, S::N1, true>;
    if (done && done[blockIdx.y]) return;
    constexpr int n = IO::n, o = IO::off, NT = C::NT, TILE = C::TILE;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* lds = reinterpret_cast<T*>(smem_raw);
    const int tid = threadIdx.x;
    const long r0 = (long)blockIdx.x * TILE;
    const Tin* src = in + blockIdx.y * bs_in + r0 * n;
    Tout* dst = out + blockIdx.y * bs_out + r0 * n;
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
template <typename Tin, typename Tout, typename T, int LOGN, int KIND>
__global__ __launch_bounds__((TileCfg<LOGN, T>::NT)) void yaxis_dst_kernel(
    const Tin* __restrict__ in, Tout* __restrict__ out, const int nx, const long bs_in, const long bs_out,
    const T scale, const T* __restrict__ tables, const int* __restrict__ done) {
    using C = TileCfg<LOGN, T>;
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
    const Tin* src = in + blockIdx.z * bs_in + poff;
    Tout* dst = out + blockIdx.z * bs_out + poff;

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

// ---------------------------------------------------------------------------
// z cyclic reduction (see dst_device.hpp for the algorithm)
// ---------------------------------------------------------------------------
template <typename T, int Q>
struct ZcrCfg {
    static constexpr int NQ = 1 << Q;
    static constexpr int G = NQ > 64 ? NQ / 64 : 1;   // lanes per system
    static constexpr int C = NQ / G;                   // rows per lane (incl. separator slot)
    static constexpr int W = 64 / G;                   // systems per wavefront
    static constexpr int LOGG = ilog2(G);
    static constexpr int LOC = G == 1 ? Q - 1 : Q - LOGG;   // levels local to a lane
    static constexpr int NT = 256;
};

// Dirichlet system (n = 2^Q - 1 rows) held as v[C] per lane (lane chunk c of
// system s, see ZcrCfg); solved in place.
template <typename T, int Q>
__device__ __forceinline__ void cr_dirichlet_lanes(T (&v)[(ZcrCfg<T, Q>::C)], const CrLevels<T, Q>& L, const int lane,
                                                   const int s, const int c) {
    using Cf = ZcrCfg<T, Q>;
    constexpr int G = Cf::G, C = Cf::C, W = Cf::W, LOC = Cf::LOC, LOGG = Cf::LOGG;
    // forward reduction inside the lane's chunk
    T toRight = T(0), toLeft = T(0);
    static_for<1, LOC + 1>([&](auto ll) {
        constexpr int l = decltype(ll)::value;
        constexpr int st = 1 << l, h = st >> 1;
        const T a = L.alpha[l];
#pragma unroll
        for (int r = st - 1; r < C - 1; r += st) v[r] += a * (v[r - h] + v[r + h]);
        if constexpr (G > 1) {
            toRight += a * v[C - 1 - h];
            toLeft += a * v[h - 1];
        }
    });

    T xleft = T(0);
    if constexpr (G == 1) {
        v[C / 2 - 1] *= L.inv_d[Q - 1];
        v[C - 1] = T(0);
    } else {
        // separator c (rows (c+1)*C-1) collects both neighbours' contributions
        const T fromRight = __shfl(toLeft, (lane + W) & 63, 64);
        const T sepv = (c < G - 1) ? v[C - 1] + toRight + fromRight : T(0);
        T sp[G];
#pragma unroll
        for (int i = 0; i < G; ++i) sp[i] = __shfl(sepv, s + i * W, 64);
        // remaining levels LOC+1..Q-1 over the G-1 separators (index i <-> row (i+1)C-1)
        static_for<1, LOGG>([&](auto ll) {
            constexpr int lp = decltype(ll)::value;
            constexpr int st = 1 << lp, h = st >> 1;
            const T a = L.alpha[LOC + lp];
#pragma unroll
            for (int i = st - 1; i < G - 1; i += st) sp[i] += a * (sp[i - h] + sp[i + h]);
        });
        sp[G / 2 - 1] *= L.inv_d[Q - 1];
        static_for<1, LOGG>([&](auto ll) {
            constexpr int lp = LOGG - decltype(ll)::value;   // LOGG-1 .. 1
            constexpr int st = 1 << lp, h = st >> 1;
            const T cc = L.c[LOC + lp - 1], id = L.inv_d[LOC + lp - 1];
#pragma unroll
            for (int i = h - 1; i < G - 1; i += st) {
                const T left = i - h >= 0 ? sp[i - h] : T(0);
                const T right = i + h < G - 1 ? sp[i + h] : T(0);
                sp[i] = (sp[i] - cc * (left + right)) * id;
            }
        });
        // this lane's boundary values
        T xl = T(0), xr = T(0);
#pragma unroll
        for (int i = 0; i < G - 1; ++i) {
            if (i == c - 1) xl = sp[i];
            if (i == c) xr = sp[i];
        }
        xleft = xl;
        v[C - 1] = xr;
    }

    // back substitution inside the chunk
    static_for<0, LOC>([&](auto ll) {
        constexpr int l = LOC - decltype(ll)::value;   // LOC .. 1
        constexpr int st = 1 << l, h = st >> 1;
        const T cc = L.c[l - 1], id = L.inv_d[l - 1];
#pragma unroll
        for (int r = h - 1; r < C - 1; r += st) {
            const T left = r - h >= 0 ? v[r - h] : xleft;
            v[r] = (v[r] - cc * (left + v[r + h])) * id;
        }
    });
}

// Ts: element type in memory; T: arithmetic type (double for FP32 plans, whose
// near-singular low modes would otherwise lose accuracy in the elimination).
template <typename Ts, typename T, int Q>
__global__ __launch_bounds__(256) void zcr_kernel(Ts* __restrict__ data, const long plane, const int nx,
                                                  const Ts* __restrict__ lmx, const Ts* __restrict__ lmy,
                                                  const T b, const T d0, const T gamma, const long bs,
                                                  const int* __restrict__ done) {
    using Cf = ZcrCfg<T, Q>;
    if (done && done[blockIdx.y]) return;
    constexpr int NQ = Cf::NQ, C = Cf::C, W = Cf::W;
    const int lane = threadIdx.x & 63;
    const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int s = lane % W, c = lane / W;
    const long m = wave * W + s;
    const bool valid = m < plane;
    const long mm = valid ? m : 0;
    Ts* base = data + blockIdx.y * bs + mm;
    const int ky = (int)(mm / nx), kx = (int)(mm - (long)ky * nx);

    T v[C];
#pragma unroll
    for (int r = 0; r < C; ++r) {
        const int row = c * C + r;
        v[r] = (valid && row < NQ - 1) ? (T)base[(long)row * plane] * gamma : T(0);
    }
    CrLevels<T, Q> L;
    cr_levels<T, Q>(L, b, d0 - (T)lmx[kx] - (T)lmy[ky]);
    cr_dirichlet_lanes<T, Q>(v, L, lane, s, c);

    if (valid) {
#pragma unroll
        for (int r = 0; r < C; ++r) {
            const int row = c * C + r;
            if (row < NQ - 1) base[(long)row * plane] = (Ts)v[r];
        }
    }
}

// ---------------------------------------------------------------------------
// Periodic z (LaplCube<..., periodic^3>): after the half-complex real DFT in
// x and y, mode (kx,ky) obeys the circulant tridiagonal system
//   b u_{i-1} + d u_i + b u_{i+1} = g_i,  i = 0..n-1 (indices mod n), n = 2^Q,
// the system the reference's z pFFT_1 + divide + pFFT solves
// (src/lapl_cube.cpp:13-30, 70-84, 124-141; pFFT(pFFT_1 x) = (n/2) x gives
// the rhs scale gamma = dz n / lz).  Cyclic reduction of a circulant keeps
// the coefficients uniform too: lane chunk c holds rows [c*C, (c+1)*C) and its
// first row is the separator.  Level l updates rows at multiples of 2^l; the
// separator's updates are additive, so each lane accumulates its own part and
// the part it owes the next chunk's separator, and one shuffle combines them.
// After Q levels row 0 obeys (d_Q + 2 c_Q) u_0 = v_0.
// Mode (0,0) is singular (d = -2b): the reference sets the zero coefficient to
// 0 after dividing by zero (src/lapl_cube.cpp:77, 82-84), i.e. it returns the
// zero-mean solution of the system with the mean of g removed.  That system is
// solved with u_0 = 0 as the Dirichlet system of rows 1..n-1 (well
// conditioned), then the mean is subtracted.
// ---------------------------------------------------------------------------
template <typename T, int Q>
struct CrLevelsP {
    T alpha[Q + 1];   // alpha[l], l = 1..Q
    T c[Q];           // c[l],  l = 0..Q-1
    T inv_d[Q];       // 1/d[l], l = 0..Q-1
    T inv_top;        // 1/(d_Q + 2 c_Q)
};

template <typename T, int Q>
__device__ __forceinline__ void cr_levels_periodic(CrLevelsP<T, Q>& L, T b, T d) {
    T cl = b, dl = d;
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        L.c[l] = cl;
        L.inv_d[l] = T(1) / dl;
        const T t = cl * cl * L.inv_d[l];
        L.alpha[l + 1] = -cl * L.inv_d[l];
        dl = dl - T(2) * t;
        cl = -t;
    }
    L.inv_top = T(1) / (dl + T(2) * cl);
}

// sum of x over the G lanes of one system (lanes s + i*W)
template <int G, int W, typename T>
__device__ __forceinline__ T system_sum(T x) {
#pragma unroll
    for (int o = W; o < 64; o <<= 1) x += __shfl_xor(x, o, 64);
    return x;
}

template <typename Ts, typename T, int Q>
__global__ __launch_bounds__(256) void zcrp_kernel(Ts* __restrict__ data, const long plane, const int nx,
                                                   const Ts* __restrict__ lmx, const Ts* __restrict__ lmy,
                                                   const T b, const T d0, const T gamma, const long bs,
                                                   const int* __restrict__ done) {
    using Cf = ZcrCfg<T, Q>;
    if (done && done[blockIdx.y]) return;
    constexpr int NQ = Cf::NQ, G = Cf::G, C = Cf::C, W = Cf::W, LOGG = Cf::LOGG;
    constexpr int LOC = Q - LOGG;   // levels local to a lane (all of them when G = 1)
    const int lane = threadIdx.x & 63;
    const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int s = lane % W, c = lane / W;
    const long m = wave * W + s;
    const bool valid = m < plane;
    const long mm = valid ? m : 0;
    Ts* base = data + blockIdx.y * bs + mm;
    const int ky = (int)(mm / nx), kx = (int)(mm - (long)ky * nx);

    T v[C];
#pragma unroll
    for (int r = 0; r < C; ++r) v[r] = valid ? (T)base[(long)(c * C + r) * plane] * gamma : T(0);

    if (mm != 0) {
        CrLevelsP<T, Q> L;
        cr_levels_periodic<T, Q>(L, b, d0 - (T)lmx[kx] - (T)lmy[ky]);
        T own = T(0), toNext = T(0);
        static_for<1, LOC + 1>([&](auto ll) {
            constexpr int l = decltype(ll)::value;
            constexpr int st = 1 << l, h = st >> 1;
            const T a = L.alpha[l];
            if constexpr (G == 1) {
                own += a * (v[h] + v[C - h]);
            } else {
                own += a * v[h];
                toNext += a * v[C - h];
            }
#pragma unroll
            for (int r = st; r < C; r += st) v[r] += a * (v[r - h] + v[r + h]);
        });
        T xright;
        if constexpr (G == 1) {
            v[0] = (v[0] + own) * L.inv_top;
            xright = v[0];
        } else {
            const T sepv = v[0] + own + __shfl(toNext, (lane - W) & 63, 64);
            T sp[G];
#pragma unroll
            for (int i = 0; i < G; ++i) sp[i] = __shfl(sepv, s + i * W, 64);
            static_for<1, LOGG + 1>([&](auto ll) {
                constexpr int lp = decltype(ll)::value;
                constexpr int st = 1 << lp, h = st >> 1;
                const T a = L.alpha[LOC + lp];
#pragma unroll
                for (int i = 0; i < G; i += st) sp[i] += a * (sp[(i - h + G) % G] + sp[(i + h) % G]);
            });
            sp[0] *= L.inv_top;
            static_for<0, LOGG>([&](auto ll) {
                constexpr int lp = LOGG - decltype(ll)::value;   // LOGG .. 1
                constexpr int st = 1 << lp, h = st >> 1;
                const T cc = L.c[LOC + lp - 1], id = L.inv_d[LOC + lp - 1];
#pragma unroll
                for (int i = h; i < G; i += st) sp[i] = (sp[i] - cc * (sp[i - h] + sp[(i + h) % G])) * id;
            });
            T x0 = T(0), xr = T(0);
#pragma unroll
            for (int i = 0; i < G; ++i) {
                if (i == c) x0 = sp[i];
                if (i == (c + 1) % G) xr = sp[i];
            }
            v[0] = x0;
            xright = xr;
        }
        static_for<0, LOC>([&](auto ll) {
            constexpr int l = LOC - decltype(ll)::value;   // LOC .. 1
            constexpr int st = 1 << l, h = st >> 1;
            const T cc = L.c[l - 1], id = L.inv_d[l - 1];
#pragma unroll
            for (int r = h; r < C; r += st) {
                const T right = r + h < C ? v[r + h] : xright;
                v[r] = (v[r] - cc * (v[r - h] + right)) * id;
            }
        });
    } else {
        // singular mode (0,0): remove the mean, pin u_0 = 0, solve rows 1..n-1
        // as a Dirichlet system, return the zero-mean solution
        T sum = T(0);
#pragma unroll
        for (int r = 0; r < C; ++r) sum += v[r];
        const T mean = system_sum<G, W>(sum) * (T(1) / T(NQ));
        // shift to Dirichlet rows 1..n-1: slot r <- row c*C + r + 1
        T nxt = T(0);
        if constexpr (G > 1) nxt = __shfl(v[0], (lane + W) & 63, 64);   // row (c+1)*C
#pragma unroll
        for (int r = 0; r + 1 < C; ++r) v[r] = v[r + 1] - mean;
        v[C - 1] = nxt - mean;
        CrLevels<T, Q> Ld;
        cr_levels<T, Q>(Ld, b, d0);
        cr_dirichlet_lanes<T, Q>(v, Ld, lane, s, c);
        T prev = T(0);
        if constexpr (G > 1) prev = __shfl(v[C - 1], (lane - W) & 63, 64);   // row c*C
#pragma unroll
        for (int r = C - 1; r > 0; --r) v[r] = v[r - 1];
        v[0] = c == 0 ? T(0) : prev;
        sum = T(0);
#pragma unroll
        for (int r = 0; r < C; ++r) sum += v[r];
        const T mu = system_sum<G, W>(sum) * (T(1) / T(NQ));
#pragma unroll
        for (int r = 0; r < C; ++r) v[r] -= mu;
    }

    if (valid) {
#pragma unroll
        for (int r = 0; r < C; ++r) base[(long)(c * C + r) * plane] = (Ts)v[r];
    }
}

// ---------------------------------------------------------------------------
// LaplRect x-direction: general-n cyclic reduction per row with per-row
// coefficients (reference: init_Mat src/lapl_rect.cpp:44-60 + gtsv :90;
// algorithm: cyclic_reduction_general src/cyclic_reduction.h:59-107).
//   D_j = -2/dx^2 - lmy[k]*s_j,  L_j = Ls_j/dx^2 (j>0),  U_j = Us_j/dx^2 (j<nx-1).
// One block per row; coefficients and rhs staged in LDS.
// ---------------------------------------------------------------------------
template <typename Ts, typename T>
__global__ __launch_bounds__(256) void rowcr_kernel(Ts* __restrict__ data, const int nx, const int q,
                                                    const Ts* __restrict__ lmy, const Ts* __restrict__ sc,
                                                    const Ts* __restrict__ Ls, const Ts* __restrict__ Us,
                                                    const double inv_dx2, const long bs,
                                                    const int* __restrict__ done) {
    if (done && done[blockIdx.y]) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* d = reinterpret_cast<T*>(smem_raw);
    T* e = d + nx;
    T* f = e + nx;
    T* bb = f + nx;
    const int k = blockIdx.x;
    Ts* row = data + blockIdx.y * bs + (long)k * nx;
    const Ts lk = lmy[k];
    for (int j = threadIdx.x; j < nx; j += blockDim.x) {
        d[j] = (T)(-2.0 * inv_dx2 - (double)(lk * sc[j]));
        e[j] = j > 0 ? (T)(Ls[j] * inv_dx2) : T(0);
        f[j] = j < nx - 1 ? (T)(Us[j] * inv_dx2) : T(0);
        bb[j] = (T)row[j];
    }
    __syncthreads();
    for (int l = 1; l < q; ++l) {
        const int st = 1 << l, h = st >> 1;
        for (int j = st - 1 + threadIdx.x * st; j < nx; j += blockDim.x * st) {
            const T alpha = -e[j] / d[j - h];
            if (j + h < nx) {
                const T gamma = -f[j] / d[j + h];
                d[j] += alpha * f[j - h] + gamma * e[j + h];
                bb[j] += alpha * bb[j - h] + gamma * bb[j + h];
                e[j] = alpha * e[j - h];
                f[j] = gamma * f[j + h];
            } else {
                d[j] += alpha * f[j - h];
                bb[j] += alpha * bb[j - h];
                e[j] = alpha * e[j - h];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        int top = (1 << (q - 1)) - 1;
        if (top > nx - 1) top = nx - 1;
        bb[top] = bb[top] / d[top];
    }
    __syncthreads();
    for (int l = q - 1; l > 0; --l) {
        const int st = 1 << l, h = st >> 1;
        for (int j = h - 1 + threadIdx.x * st; j < nx; j += blockDim.x * st) {
            T acc = bb[j];
            if (j - h >= 0) acc -= e[j] * bb[j - h];
            if (j + h < nx) acc -= f[j] * bb[j + h];
            bb[j] = acc / d[j];
        }
        __syncthreads();
    }
    for (int j = threadIdx.x; j < nx; j += blockDim.x) row[j] = (Ts)bb[j];
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
namespace {
template <typename K>
hipError_t allow_lds(K kernel, size_t bytes) {
    if (bytes <= 65536) return hipSuccess;
    return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes);
}
}  // namespace

template <typename Tin, typename Tout, typename T, int LOGN, bool ACCUM, int KIND>
static hipError_t plane_k(const PlaneArgs& a) {
    using C = PlaneCfg<LOGN>;
    auto k = plane_dst_kernel<Tin, Tout, T, LOGN, ACCUM, KIND>;
    const size_t lds = C::template lds_bytes<T>();
    hipError_t e = allow_lds(k, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(a.nplanes, a.batch), dim3(C::NT), lds, a.stream, (const Tin*)a.in, (Tout*)a.out,
                       a.bs_in, a.bs_out, (T)a.sx, (T)a.sy, (const T*)a.tables, a.done);
    return hipGetLastError();
}

template <typename Tin, typename Tout, typename T, int LOGN>
static hipError_t plane_t(const PlaneArgs& a) {
    if (a.kind == kDst)
        return a.accum ? plane_k<Tin, Tout, T, LOGN, true, kDst>(a) : plane_k<Tin, Tout, T, LOGN, false, kDst>(a);
    if constexpr (std::is_same_v<Tin, T> && std::is_same_v<Tout, T>) {
        if (a.accum) return hipErrorInvalidValue;
        if (a.kind == kRdftFwd) return plane_k<Tin, Tout, T, LOGN, false, kRdftFwd>(a);
        if (a.kind == kRdftInv) return plane_k<Tin, Tout, T, LOGN, false, kRdftInv>(a);
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
    hipLaunchKernelGGL(k, dim3(per * kXcds, a.batch), dim3(C::NT), lds, a.stream, a.f, (const Tu*)a.u, a.out,
                       a.partial, a.bs, a.nplanes, a.ix, a.iy, a.iz, (float)a.sx, (float)a.sy,
                       (const float*)a.tables, a.done);
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

template <typename Tin, typename Tout, typename T, int LOGN, bool ACCUM, int KIND>
static hipError_t xaxis_k(const AxisArgs& a) {
    using C = TileCfg<LOGN, T>;
    auto k = xaxis_dst_kernel<Tin, Tout, T, LOGN, ACCUM, KIND>;
    const size_t lds = C::lds_x();
    hipError_t e = allow_lds(k, lds);
    if (e != hipSuccess) return e;
    const long tiles = (a.nrows + C::TILE - 1) / C::TILE;
    hipLaunchKernelGGL(k, dim3((unsigned)tiles, a.batch), dim3(C::NT), lds, a.stream, (const Tin*)a.in,
                       (Tout*)a.out, a.nrows, a.bs_in, a.bs_out, (T)a.scale, (const T*)a.tables, a.done);
    return hipGetLastError();
}

template <typename Tin, typename Tout, typename T, int LOGN>
static hipError_t xaxis_t(const AxisArgs& a) {
    if (a.kind == kDst)
        return a.accum ? xaxis_k<Tin, Tout, T, LOGN, true, kDst>(a) : xaxis_k<Tin, Tout, T, LOGN, false, kDst>(a);
    if constexpr (std::is_same_v<Tin, T> && std::is_same_v<Tout, T>) {
        if (a.accum) return hipErrorInvalidValue;
        if (a.kind == kRdftFwd) return xaxis_k<Tin, Tout, T, LOGN, false, kRdftFwd>(a);
        if (a.kind == kRdftInv) return xaxis_k<Tin, Tout, T, LOGN, false, kRdftInv>(a);
    }
    return hipErrorInvalidValue;
}

template <typename Tin, typename Tout, typename T, int LOGN, int KIND>
static hipError_t yaxis_k(const AxisArgs& a) {
    using C = TileCfg<LOGN, T>;
    auto k = yaxis_dst_kernel<Tin, Tout, T, LOGN, KIND>;
    const size_t lds = C::lds_y();
    hipError_t e = allow_lds(k, lds);
    if (e != hipSuccess) return e;
    const int tiles = (a.nx + C::TILE - 1) / C::TILE;
    hipLaunchKernelGGL(k, dim3(tiles, a.nplanes, a.batch), dim3(C::NT), lds, a.stream, (const Tin*)a.in,
                       (Tout*)a.out, a.nx, a.bs_in, a.bs_out, (T)a.scale, (const T*)a.tables, a.done);
    return hipGetLastError();
}

template <typename Tin, typename Tout, typename T, int LOGN>
static hipError_t yaxis_t(const AxisArgs& a) {
    if (a.kind == kDst) return yaxis_k<Tin, Tout, T, LOGN, kDst>(a);
    if constexpr (std::is_same_v<Tin, T> && std::is_same_v<Tout, T>) {
        if (a.kind == kRdftFwd) return yaxis_k<Tin, Tout, T, LOGN, kRdftFwd>(a);
        if (a.kind == kRdftInv) return yaxis_k<Tin, Tout, T, LOGN, kRdftInv>(a);
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

template <typename Ts, typename T, int Q>
static hipError_t zcr_k(const ZcrArgs& a) {
    using C = ZcrCfg<T, Q>;
    const long lanes = ((a.plane + C::W - 1) / C::W) * 64;
    const unsigned blocks = (unsigned)((lanes + C::NT - 1) / C::NT);
    auto k = a.periodic ? zcrp_kernel<Ts, T, Q> : zcr_kernel<Ts, T, Q>;
    hipLaunchKernelGGL(k, dim3(blocks, a.batch), dim3(C::NT), 0, a.stream, (Ts*)a.data, a.plane, a.nx,
                       (const Ts*)a.lmx, (const Ts*)a.lmy, (T)a.b, (T)a.d0, (T)a.gamma, a.bs, a.done);
    return hipGetLastError();
}

template <typename T, int Q>
static hipError_t zcr_t(const ZcrArgs& a) {
    if constexpr (std::is_same_v<T, float>) {
        if (a.wide) return zcr_k<float, double, Q>(a);
    }
    return zcr_k<T, T, Q>(a);
}

template <typename T>
hipError_t launch_zcr(int q, const ZcrArgs& a) {
    switch (q) {
        case 1: return zcr_t<T, 1>(a);
        case 2: return zcr_t<T, 2>(a);
        case 3: return zcr_t<T, 3>(a);
        case 4: return zcr_t<T, 4>(a);
        case 5: return zcr_t<T, 5>(a);
        case 6: return zcr_t<T, 6>(a);
        case 7: return zcr_t<T, 7>(a);
        case 8: return zcr_t<T, 8>(a);
        case 9: return zcr_t<T, 9>(a);
        default: return hipErrorInvalidValue;
    }
}

template <typename Ts, typename T>
static hipError_t rowcr_k(const RowCrArgs& a) {
    int q = 0;
    while ((1 << q) < a.nx + 1) q++;
    const size_t lds = 4 * sizeof(T) * (size_t)a.nx;
    hipError_t e = allow_lds(rowcr_kernel<Ts, T>, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((rowcr_kernel<Ts, T>), dim3(a.nrows, a.batch), dim3(256), lds, a.stream, (Ts*)a.data, a.nx,
                       q, (const Ts*)a.lmy, (const Ts*)a.sc, (const Ts*)a.Ls, (const Ts*)a.Us, a.inv_dx2, a.bs,
                       a.done);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_rowcr(const RowCrArgs& a) {
    if constexpr (std::is_same_v<T, float>) {
        if (a.wide) return rowcr_k<float, double>(a);
    }
    return rowcr_k<T, T>(a);
}

template hipError_t launch_plane<float, float, float>(int, const PlaneArgs&);
template hipError_t launch_plane<double, double, double>(int, const PlaneArgs&);
template hipError_t launch_plane<double, float, float>(int, const PlaneArgs&);
template hipError_t launch_plane<float, double, float>(int, const PlaneArgs&);
template bool plane_fits<float>(int);
template bool plane_fits<double>(int);
template hipError_t launch_xaxis<float, float, float>(int, const AxisArgs&);
template hipError_t launch_xaxis<double, double, double>(int, const AxisArgs&);
template hipError_t launch_xaxis<double, float, float>(int, const AxisArgs&);
template hipError_t launch_xaxis<float, double, float>(int, const AxisArgs&);
template hipError_t launch_yaxis<float, float, float>(int, const AxisArgs&);
template hipError_t launch_yaxis<double, double, double>(int, const AxisArgs&);
template hipError_t launch_yaxis<double, float, float>(int, const AxisArgs&);
template hipError_t launch_yaxis<float, double, float>(int, const AxisArgs&);
template hipError_t launch_zcr<float>(int, const ZcrArgs&);
template hipError_t launch_zcr<double>(int, const ZcrArgs&);
template hipError_t launch_rowcr<float>(const RowCrArgs&);
template hipError_t launch_rowcr<double>(const RowCrArgs&);

}  // namespace fdmgpu
