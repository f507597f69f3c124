// Tridiagonal kernels (gfx950): constant-coefficient cyclic reduction along
// z for every (kx,ky) mode (Dirichlet and circulant), and LaplRect's
// per-row cyclic reduction along x.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "dst_device.hpp"
#include "kernel_util.hpp"
#include "kernels.hpp"

namespace fdmgpu {

constexpr int kXcdSlabs = 8;   // XCDs of MI355X; workgroups are dealt to them round-robin

// 1-D grid of per * kXcdSlabs blocks covering `blocks` blocks of each of
// `batch` problems: the logical block (problem, block) is dealt out in
// contiguous slabs per XCD so neighbouring blocks of a problem, whose row
// segments share partial cache lines, run under one L2.  False for the
// padding blocks.
__device__ __forceinline__ bool xcd_slab_block(const long blocks, const int batch, int& p, long& blk) {
    const long total = blocks * batch;
    const long per = (total + kXcdSlabs - 1) / kXcdSlabs;
    const long logical = (blockIdx.x % kXcdSlabs) * per + blockIdx.x / kXcdSlabs;
    if (logical >= total) return false;
    p = (int)(logical / blocks);
    blk = logical - (long)p * blocks;
    return true;
}

static unsigned xcd_slab_grid(const long blocks, const int batch) {
    const long total = blocks * batch;
    return (unsigned)((total + kXcdSlabs - 1) / kXcdSlabs * kXcdSlabs);
}

// ---------------------------------------------------------------------------
// z cyclic reduction (see dst_device.hpp for the algorithm)
// ---------------------------------------------------------------------------
// GM multiplies the lanes per system (a tunable: more lanes = more
// parallelism and fewer registers per lane, more shuffles).
template <typename T, int Q, int GM = 1>
struct ZcrCfg {
    static constexpr int NQ = 1 << Q;
    static constexpr int G = (NQ > 64 ? NQ / 64 : 1) * GM;   // lanes per system
    static constexpr int C = NQ / G;                   // rows per lane (incl. separator slot)
    static constexpr int W = 64 / G;                   // systems per wavefront
    static constexpr int LOGG = ilog2(G);
    static constexpr int LOC = G == 1 ? Q - 1 : Q - LOGG;   // levels local to a lane
    static constexpr int NT = 256;
};

// Dirichlet system (n = 2^Q - 1 rows) held as v[C] per lane (lane chunk c of
// system s, see ZcrCfg); solved in place.
template <typename T, int Q, int GM>
__device__ __forceinline__ void cr_dirichlet_lanes(T (&v)[(ZcrCfg<T, Q, GM>::C)], const CrLevels<T, Q>& L,
                                                   const int lane, const int s, const int c) {
    using Cf = ZcrCfg<T, Q, GM>;
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
template <typename Ts, typename T, int Q, int GM>
__global__ __launch_bounds__(256) void zcr_kernel(const BatchPtr<Ts> data, const long plane, const int nx,
                                                  const Ts* __restrict__ lmx, const Ts* __restrict__ lmy,
                                                  const T b, const T d0, const T gamma,
                                                  const int* __restrict__ done, const long blocks,
                                                  const int batch) {
    using Cf = ZcrCfg<T, Q, GM>;
    int p;
    long blk;
    if (!xcd_slab_block(blocks, batch, p, blk) || (done && done[p])) return;
    constexpr int NQ = Cf::NQ, C = Cf::C, W = Cf::W;
    const int lane = threadIdx.x & 63;
    const long wave = (blk * blockDim.x + threadIdx.x) >> 6;
    const int s = lane % W, c = lane / W;
    const long m = wave * W + s;
    const bool valid = m < plane;
    const long mm = valid ? m : 0;
    const auto base = as_global(data(p) + mm);
    const int ky = (int)(mm / nx), kx = (int)(mm - (long)ky * nx);

    T v[C];
#pragma unroll
    for (int r = 0; r < C; ++r) {
        const int row = c * C + r;
        // branch-free: the row is clamped into the field and the value
        // selected, so all C loads are in flight together
        const T x = (T)base[(long)(row < NQ - 1 ? row : NQ - 2) * plane];
        v[r] = (valid && row < NQ - 1) ? x * gamma : T(0);
    }
    CrLevels<T, Q> L;
    cr_levels<T, Q>(L, b, d0 - (T)lmx[kx] - (T)lmy[ky]);
    cr_dirichlet_lanes<T, Q, GM>(v, L, lane, s, c);

    if (valid) {
#pragma unroll
        for (int r = 0; r < C; ++r) {
            const int row = c * C + r;
            if (row < NQ - 1) base[(long)row * plane] = (Ts)v[r];
        }
    }
}

// ---------------------------------------------------------------------------
// Strip layout (Dirichlet): a wavefront holds one 32-row chunk of 64
// consecutive systems, lane = system, so every load and store instruction of
// a wave covers 64 consecutive modes of one z-plane (256 B in FP32, 512 B in
// FP64) instead of the lane layout's G short runs.  The G = 2^Q / 32 chunk
// waves of a strip exchange their separator contributions through LDS (one
// barrier).  S strips per block.  The arithmetic is the lane layout's: the
// same levels in the same order, so results are identical.
// Blocks go round-robin over the 8 XCDs; the logical block (problem, strip
// group) is dealt out in contiguous slabs per XCD, so neighbouring blocks,
// whose row segments share partial cache lines, run under one L2.
// ---------------------------------------------------------------------------
template <int Q, int CR = 32>
struct ZcrStripCfg {
    static constexpr int NQ = 1 << Q;
    static constexpr int C = NQ >= CR ? CR : NQ;   // rows per lane (incl. separator slot)
    static constexpr int G = NQ / C;               // chunk waves per strip
    static constexpr int LOGG = ilog2(G);
    static constexpr int LOC = G == 1 ? Q - 1 : Q - LOGG;
};

template <typename Ts, typename T, int Q, int S, int CR>
__global__ __launch_bounds__((64 * ZcrStripCfg<Q, CR>::G * S)) void zcr_strip_kernel(
    const BatchPtr<Ts> data, const long plane, const int nx, const Ts* __restrict__ lmx,
    const Ts* __restrict__ lmy, const T b, const T d0, const T gamma, const int* __restrict__ done,
    const int blocks, const int batch) {
    using Cf = ZcrStripCfg<Q, CR>;
    constexpr int NQ = Cf::NQ, C = Cf::C, G = Cf::G, LOGG = Cf::LOGG, LOC = Cf::LOC;
    constexpr int SYS = 64 * S;   // systems per block
    __shared__ T xch[G > 1 ? 2 * G * SYS : 1];
    int p;
    long blk;
    if (!xcd_slab_block(blocks, batch, p, blk) || (done && done[p])) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int c = w % G, strip = w / G;
    const int sys = strip * 64 + lane;
    const long m = (long)blk * SYS + sys;
    const bool valid = m < plane;
    const long mm = valid ? m : 0;
    const auto base = as_global(data(p) + mm);
    const int ky = (int)(mm / nx), kx = (int)(mm - (long)ky * nx);

    T v[C];
#pragma unroll
    for (int r = 0; r < C; ++r) {
        const int row = c * C + r;
        // branch-free: the row is clamped into the field and the value
        // selected, so all C loads are in flight together
        const T x = (T)base[(long)(row < NQ - 1 ? row : NQ - 2) * plane];
        v[r] = (valid && row < NQ - 1) ? x * gamma : T(0);
    }
    CrLevels<T, Q> L;
    cr_levels<T, Q>(L, b, d0 - (T)lmx[kx] - (T)lmy[ky]);

    // forward reduction inside the chunk
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
        // separator c (row (c+1)*C-1): this chunk's part and the next chunk's toLeft
        T* tl = xch;
        T* sv = xch + G * SYS;
        tl[c * SYS + sys] = toLeft;
        sv[c * SYS + sys] = (c < G - 1) ? v[C - 1] + toRight : T(0);
        __syncthreads();
        T sp[G];
#pragma unroll
        for (int i = 0; i < G; ++i) sp[i] = (i < G - 1) ? sv[i * SYS + sys] + tl[(i + 1) * SYS + sys] : T(0);
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

template <typename Ts, typename T, int Q, int GM>
__global__ __launch_bounds__(256) void zcrp_kernel(const BatchPtr<Ts> data, const long plane, const int nx,
                                                   const Ts* __restrict__ lmx, const Ts* __restrict__ lmy,
                                                   const T b, const T d0, const T gamma,
                                                   const int* __restrict__ done, const long blocks,
                                                   const int batch) {
    using Cf = ZcrCfg<T, Q, GM>;
    int p;
    long blk;
    if (!xcd_slab_block(blocks, batch, p, blk) || (done && done[p])) return;
    constexpr int NQ = Cf::NQ, G = Cf::G, C = Cf::C, W = Cf::W, LOGG = Cf::LOGG;
    constexpr int LOC = Q - LOGG;   // levels local to a lane (all of them when G = 1)
    const int lane = threadIdx.x & 63;
    const long wave = (blk * blockDim.x + threadIdx.x) >> 6;
    const int s = lane % W, c = lane / W;
    const long m = wave * W + s;
    const bool valid = m < plane;
    const long mm = valid ? m : 0;
    const auto base = as_global(data(p) + mm);
    const int ky = (int)(mm / nx), kx = (int)(mm - (long)ky * nx);

    T v[C];
#pragma unroll
    for (int r = 0; r < C; ++r) {
        const T x = (T)base[(long)(c * C + r) * plane];   // in the field for every lane (mm is clamped)
        v[r] = valid ? x * gamma : T(0);
    }

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
        cr_dirichlet_lanes<T, Q, GM>(v, Ld, lane, s, c);
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
// LaplCyl3FFT2's r systems (src/lapl_cyl.cpp:55-73) are the same family with a
// second eigenvalue: row k = i*rdiv + kz, D_j = -2/dr^2 - lm_phi[i]/r_j^2 -
// lm_z[kz] (lmy = lm_phi, s_j = 1/r_j^2, lmz = lm_z).
// One block per row; coefficients and rhs staged in LDS.
// ---------------------------------------------------------------------------
template <typename Ts, typename T>
__global__ __launch_bounds__(256) void rowcr_kernel(const BatchPtr<Ts> data, const int nx, const int q,
                                                    const Ts* __restrict__ lmy, const Ts* __restrict__ sc,
                                                    const Ts* __restrict__ Ls, const Ts* __restrict__ Us,
                                                    const double inv_dx2, const int* __restrict__ done,
                                                    const Ts* __restrict__ lmz, const int rdiv) {
    if (done && done[blockIdx.y]) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* d = reinterpret_cast<T*>(smem_raw);
    T* e = d + nx;
    T* f = e + nx;
    T* bb = f + nx;
    const int k = blockIdx.x;
    const auto row = as_global(data(blockIdx.y) + (long)k * nx);
    const Ts lk = lmy[lmz ? k / rdiv : k];
    const double lz = lmz ? (double)lmz[k % rdiv] : 0.0;
    for (int j = threadIdx.x; j < nx; j += blockDim.x) {
        d[j] = lmz ? (T)(-2.0 * inv_dx2 - (double)(lk * sc[j]) - lz) : (T)(-2.0 * inv_dx2 - (double)(lk * sc[j]));
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
// LaplRect rows that are not diagonally dominant (lm_y_scale / L_scale /
// U_scale chosen so): the reference solves every row with LAPACK ?gtsv
// (src/lapl_rect.cpp:90, src/blas.h:67-75), Gaussian elimination with partial
// pivoting -- a row interchange when |d_i| < |dl_i| -- and back substitution
// over the second superdiagonal the interchanges fill in (LAPACK's published
// algorithm).  Elimination is sequential along a row: one block per row, the
// coefficients and the right-hand side staged in LDS by all threads, the
// elimination by one thread.  Row systems are those of rowcr_kernel:
// dl_j = L_scale(j+1)/dx^2 (row j+1's sub-diagonal), d_j, du_j = U_scale(j)/dx^2.
// Singular rows (a zero pivot, where the reference aborts) are rejected on the
// host when the scales are set.
// ---------------------------------------------------------------------------
template <typename Ts, typename T>
__global__ __launch_bounds__(64) void rowgtsv_kernel(const BatchPtr<Ts> data, const int nx,
                                                     const Ts* __restrict__ lmy, const Ts* __restrict__ sc,
                                                     const Ts* __restrict__ Ls, const Ts* __restrict__ Us,
                                                     const double inv_dx2, const int* __restrict__ done,
                                                     const Ts* __restrict__ lmz, const int rdiv) {
    if (done && done[blockIdx.y]) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* dl = reinterpret_cast<T*>(smem_raw);
    T* d = dl + nx;
    T* du = d + nx;
    T* b = du + nx;
    const int k = blockIdx.x;
    const auto row = as_global(data(blockIdx.y) + (long)k * nx);
    const Ts lk = lmy[lmz ? k / rdiv : k];
    const double lz = lmz ? (double)lmz[k % rdiv] : 0.0;
    for (int j = threadIdx.x; j < nx; j += blockDim.x) {
        d[j] = (T)(-2.0 * inv_dx2 - (double)(lk * sc[j]) - lz);
        dl[j] = j + 1 < nx ? (T)(Ls[j + 1] * inv_dx2) : T(0);
        du[j] = j + 1 < nx ? (T)(Us[j] * inv_dx2) : T(0);
        b[j] = (T)row[j];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int n = nx;
        for (int i = 0; i < n - 1; i++) {
            if (fabs(d[i]) >= fabs(dl[i])) {   // no interchange
                const T fact = dl[i] / d[i];
                d[i + 1] -= fact * du[i];
                b[i + 1] -= fact * b[i];
                if (i < n - 2) dl[i] = T(0);
            } else {                           // interchange rows i and i+1
                const T fact = d[i] / dl[i];
                d[i] = dl[i];
                T temp = d[i + 1];
                d[i + 1] = du[i] - fact * temp;
                if (i < n - 2) {
                    dl[i] = du[i + 1];         // second superdiagonal
                    du[i + 1] = -fact * dl[i];
                }
                du[i] = temp;
                temp = b[i];
                b[i] = b[i + 1];
                b[i + 1] = temp - fact * b[i + 1];
            }
        }
        b[n - 1] /= d[n - 1];
        if (n > 1) b[n - 2] = (b[n - 2] - du[n - 2] * b[n - 1]) / d[n - 2];
        for (int i = n - 3; i >= 0; i--) b[i] = (b[i] - du[i] * b[i + 1] - dl[i] * b[i + 2]) / d[i];
    }
    __syncthreads();
    for (int j = threadIdx.x; j < nx; j += blockDim.x) row[j] = (Ts)b[j];
}

// Lane multipliers a z system of 2^q rows supports: 1 always, 2 when every
// lane still holds at least 2 rows; the strip layouts (3: 64 systems per
// block, 4: 128) from 32 rows up (Dirichlet only), 4 up to 128 rows (512
// threads, registers stay within budget in FP64); 5: 16-row chunks (half the
// registers per lane), 16 to 256 rows.
bool zcr_gmul_ok(int q, int gm) {
    if (gm == 1) return true;
    if (gm == 3) return q >= 5 && q <= 9;
    if (gm == 4) return q >= 5 && q <= 7;
    if (gm == 5) return q >= 4 && q <= 8;
    if (gm != 2 || q < 2) return false;
    const int NQ = 1 << q;
    const int G = (NQ > 64 ? NQ / 64 : 1) * gm;
    return NQ / G >= 2 && G <= 64;
}

template <typename Ts, typename T, int Q, int GM>
static hipError_t zcr_k(const ZcrArgs& a) {
    using C = ZcrCfg<T, Q, GM>;
    const long lanes = ((a.plane + C::W - 1) / C::W) * 64;
    const long blocks = (lanes + C::NT - 1) / C::NT;
    auto k = a.periodic ? zcrp_kernel<Ts, T, Q, GM> : zcr_kernel<Ts, T, Q, GM>;
    const BatchPtr<Ts> data{(Ts*)a.data, a.data_tab, a.bs};
    hipLaunchKernelGGL(k, dim3(xcd_slab_grid(blocks, a.batch)), dim3(C::NT), 0, a.stream, data, a.plane, a.nx,
                       (const Ts*)a.lmx, (const Ts*)a.lmy, (T)a.b, (T)a.d0, (T)a.gamma, a.done, blocks, a.batch);
    return hipGetLastError();
}

template <typename Ts, typename T, int Q, int S, int CR>
static hipError_t zcr_strip_k(const ZcrArgs& a) {
    constexpr int SYS = 64 * S;
    const long blocks = (a.plane + SYS - 1) / SYS;
    const BatchPtr<Ts> data{(Ts*)a.data, a.data_tab, a.bs};
    hipLaunchKernelGGL((zcr_strip_kernel<Ts, T, Q, S, CR>), dim3(xcd_slab_grid(blocks, a.batch)),
                       dim3(64 * ZcrStripCfg<Q, CR>::G * S), 0, a.stream, data, a.plane, a.nx, (const Ts*)a.lmx,
                       (const Ts*)a.lmy, (T)a.b, (T)a.d0, (T)a.gamma, a.done, blocks, a.batch);
    return hipGetLastError();
}

template <typename Ts, typename T, int Q>
static hipError_t zcr_g(const ZcrArgs& a) {
    // default: Dirichlet systems in the strip layout (64 consecutive modes per
    // wave) with 16-row chunks up to 128 rows (measured fastest at 128 rows:
    // 8 waves per SIMD in FP32) and 32-row chunks from 256 rows (measured 2-3%
    // faster than 16 at 256 rows, FP64 and FP32); circulant ones with 2 lanes
    // per 64 rows from 64 rows up (fewer registers per lane, twice the
    // occupancy of 1)
    const int gm = a.gmul                                             ? a.gmul
                   : (!a.periodic && Q <= 7 && zcr_gmul_ok(Q, 5)) ? 5
                   : (!a.periodic && zcr_gmul_ok(Q, 3))           ? 3
                   : (!a.periodic && zcr_gmul_ok(Q, 5))           ? 5
                   : (Q >= 6 ? 2 : 1);
    if (!zcr_gmul_ok(Q, gm) || (gm >= 3 && a.periodic)) return hipErrorInvalidValue;
    if (gm == 3) return zcr_strip_k<Ts, T, Q, 1, 32>(a);
    if constexpr (ZcrStripCfg<Q>::G <= 4) {
        if (gm == 4) return zcr_strip_k<Ts, T, Q, 2, 32>(a);
    }
    if constexpr (ZcrStripCfg<Q, 16>::G <= 16) {
        if (gm == 5) return zcr_strip_k<Ts, T, Q, 1, 16>(a);
    }
    if (gm == 1) return zcr_k<Ts, T, Q, 1>(a);
    if constexpr (Q >= 2) {
        if (gm == 2) return zcr_k<Ts, T, Q, 2>(a);
    }
    return hipErrorInvalidValue;
}

template <typename T, int Q>
static hipError_t zcr_t(const ZcrArgs& a) {
    if constexpr (std::is_same_v<T, float>) {
        if (a.wide) return zcr_g<float, double, Q>(a);
    }
    return zcr_g<T, T, Q>(a);
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
    const BatchPtr<Ts> data{(Ts*)a.data, a.data_tab, a.bs};
    hipLaunchKernelGGL((rowcr_kernel<Ts, T>), dim3(a.nrows, a.batch), dim3(256), lds, a.stream, data, a.nx, q,
                       (const Ts*)a.lmy, (const Ts*)a.sc, (const Ts*)a.Ls, (const Ts*)a.Us, a.inv_dx2, a.done,
                       (const Ts*)a.lmz, a.rdiv);
    return hipGetLastError();
}

template <typename Ts, typename T>
static hipError_t rowgtsv_k(const RowCrArgs& a) {
    const size_t lds = 4 * sizeof(T) * (size_t)a.nx;
    hipError_t e = allow_lds(rowgtsv_kernel<Ts, T>, lds);
    if (e != hipSuccess) return e;
    const BatchPtr<Ts> data{(Ts*)a.data, a.data_tab, a.bs};
    hipLaunchKernelGGL((rowgtsv_kernel<Ts, T>), dim3(a.nrows, a.batch), dim3(64), lds, a.stream, data, a.nx,
                       (const Ts*)a.lmy, (const Ts*)a.sc, (const Ts*)a.Ls, (const Ts*)a.Us, a.inv_dx2, a.done,
                       (const Ts*)a.lmz, a.rdiv);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_rowcr(const RowCrArgs& a) {
    if (a.pivot) {
        if constexpr (std::is_same_v<T, float>) {
            if (a.wide) return rowgtsv_k<float, double>(a);
        }
        return rowgtsv_k<T, T>(a);
    }
    if constexpr (std::is_same_v<T, float>) {
        if (a.wide) return rowcr_k<float, double>(a);
    }
    return rowcr_k<T, T>(a);
}

// ---------------------------------------------------------------------------
// LaplRectFFT2 (src/lapl_rect.cpp:162-172): after the transforms along y and
// x every coefficient is divided by -(lm_y[k] lm_y_scale[j] + lm_x[j]), in T
// as the reference does; the double-periodic plan sets the (0,0) coefficient
// (0/0) to 1.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void rect2_divide_kernel(const BatchPtr<T> data, const int nx, const int ny,
                                                          const T* __restrict__ lmy, const T* __restrict__ lmx,
                                                          const T* __restrict__ sc, const bool zero_mode,
                                                          const int* __restrict__ done) {
    if (done && done[blockIdx.y]) return;
    const long n = (long)nx * ny;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = (int)(i / nx), j = (int)(i - (long)k * nx);
    const auto d = as_global(data(blockIdx.y));
    T v = d[i];
    v /= -lmy[k] * sc[j] - lmx[j];
    if (zero_mode && i == 0) v = T(1);
    d[i] = v;
}

template <typename T>
hipError_t launch_rect2_divide(const Rect2DivArgs& a) {
    const long n = (long)a.nx * a.ny;
    const BatchPtr<T> data{(T*)a.data, a.data_tab, a.bs};
    hipLaunchKernelGGL(rect2_divide_kernel<T>, dim3((unsigned)((n + 255) / 256), a.batch), dim3(256), 0, a.stream,
                       data, a.nx, a.ny, (const T*)a.lmy, (const T*)a.lmx, (const T*)a.sc, a.zero_mode, a.done);
    return hipGetLastError();
}

template hipError_t launch_rect2_divide<float>(const Rect2DivArgs&);
template hipError_t launch_rect2_divide<double>(const Rect2DivArgs&);
template hipError_t launch_zcr<float>(int, const ZcrArgs&);
template hipError_t launch_zcr<double>(int, const ZcrArgs&);
template hipError_t launch_rowcr<float>(const RowCrArgs&);
template hipError_t launch_rowcr<double>(const RowCrArgs&);

}  // namespace fdmgpu
