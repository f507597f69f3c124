// FP64 side of FP32+refinement: residual stencil, norms, convergence flags.
//
// The operator is the one the reference's solve inverts (the 7-point /
// 5-point discrete Laplacian with zero Dirichlet boundary, spacings as the
// reference's eigenvalue tables use them, src/lapl_cube.cpp:145-172,
// src/lapl_rect.cpp:13-22, 44-60):
//   (A u)_i = (u_{x-1} - 2u + u_{x+1}) * cx_i  (+ LaplRect scales on x)
//           + (u_{y-1} - 2u + u_{y+1}) * iy * sy_x
//           + (u_{z-1} - 2u + u_{z+1}) * iz          (cube only)
// Norm partials are reduced in a fixed order (no atomics), so a problem's
// residual and iteration count do not depend on what else is in the batch.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "refine.hpp"

namespace fdmgpu {

constexpr int kResBlock = 256;

template <typename Tr>
__device__ __forceinline__ void block_sum2(double& a, double& b, double* sh) {
    const int tid = threadIdx.x;
    sh[tid] = a;
    sh[kResBlock + tid] = b;
    __syncthreads();
#pragma unroll
    for (int s = kResBlock / 2; s > 0; s >>= 1) {
        if (tid < s) {
            sh[tid] += sh[tid + s];
            sh[kResBlock + tid] += sh[kResBlock + tid + s];
        }
        __syncthreads();
    }
    a = sh[0];
    b = sh[kResBlock];
}

// r = f - A u (FP64), r32 = (float) r; partial[p][c] = (sum r^2, sum f^2)
// over chunk c of problem blockIdx.y.  A chunk is a contiguous range of rows
// (row = z * ny + y), one wave per row and the lanes along x.  The gridDim.x
// chunks (a multiple of 8) are dealt to the 8 XCDs round-robin, so
// chunk = (bid % 8) * (nblk / 8) + bid / 8 gives every XCD one contiguous slab
// of z-planes and the u planes at z +- 1 are re-read from that XCD's L2.
// A launch for already-converged problems costs a few microseconds.
template <bool THREE, bool SCALED>
__global__ __launch_bounds__(kResBlock) void residual_kernel(
    const BatchPtr<const double> f, const BatchPtr<const double> u, float* __restrict__ r32,
    double* __restrict__ partial, const int nx, const int ny, const int nz, const double ix, const double iy,
    const double iz, const double* __restrict__ sy, const double* __restrict__ lx,
    const double* __restrict__ ux, const long bs, const int* __restrict__ done) {
    __shared__ double sh[2 * kResBlock];
    const int p = blockIdx.y;
    if (done && done[p]) return;
    const auto fp = as_global(f(p));
    const auto up = as_global(u(p));
    float* rp = r32 ? r32 + p * bs : nullptr;
    const long pl = (long)nx * ny;
    const int nblk = gridDim.x;
    const int c = (blockIdx.x % 8) * (nblk / 8) + blockIdx.x / 8;
    const int nrows = ny * nz;
    const int chunk = (nrows + nblk - 1) / nblk;
    const int row_end = min(nrows, (c + 1) * chunk);
    const int lane = threadIdx.x & 63;
    double rr = 0.0, ff = 0.0;
    for (int row = c * chunk + (int)(threadIdx.x >> 6); row < row_end; row += kResBlock / 64) {
        const int y = row % ny;
        const int z = row / ny;
        const long base = (long)row * nx;
        for (int x = lane; x < nx; x += 64) {
            const long idx = base + x;
            const double uc = up[idx];
            const double xm = x > 0 ? up[idx - 1] : 0.0, xp = x < nx - 1 ? up[idx + 1] : 0.0;
            const double ym = y > 0 ? up[idx - nx] : 0.0, yp = y < ny - 1 ? up[idx + nx] : 0.0;
            double au;
            if constexpr (SCALED) {
                au = (lx[x] * xm - 2.0 * uc + ux[x] * xp) * ix + (ym - 2.0 * uc + yp) * iy * sy[x];
            } else {
                au = (xm - 2.0 * uc + xp) * ix + (ym - 2.0 * uc + yp) * iy;
            }
            if constexpr (THREE) {
                const double zm = z > 0 ? up[idx - pl] : 0.0, zp = z < nz - 1 ? up[idx + pl] : 0.0;
                au += (zm - 2.0 * uc + zp) * iz;
            }
            const double fv = fp[idx];
            const double r = fv - au;
            if (r32) rp[idx] = (float)r;
            rr += r * r;
            ff += fv * fv;
        }
    }
    block_sum2<double>(rr, ff, sh);
    if (threadIdx.x == 0) {
        partial[((long)p * nblk + c) * 2 + 0] = rr;
        partial[((long)p * nblk + c) * 2 + 1] = ff;
    }
}

// 3-D residual: a block of TX x TY threads owns the tile x in [x0, x0+TX),
// y in [y0, y0+TY) of one z-chunk and marches through it.  The u plane of the
// current z sits in LDS (two buffers, one barrier per plane) with a one-point
// halo, so the x / y neighbours are LDS reads; u at z-1, z, z+1 stay in
// registers.  Loads run two planes ahead of use: while plane z is computed,
// u at z+2 and z+3 and f and the halo values at z+1 and z+2 are in flight.  Each u value is read from
// HBM once per chunk plus the chunk's two boundary planes and the tile's halo
// rows.  Tiles are as wide as the rows where possible (no x halo).
// Blocks go round-robin over the 8 XCDs; the logical block index (problem,
// chunk, tile) is dealt out in contiguous slabs per XCD, so neighbouring
// tiles, whose halo rows overlap, share an L2.
// Norm partials: one per (z-granule of kZGran planes, tile, wave), written by
// the wave that covers it: partial[p][(g * tiles + tile) * NW + w].  The
// chunk length (a whole number of granules, chosen per launch to fill the
// GPU) changes neither the layout nor the summation order.
constexpr int kZGran = 8;
constexpr int kXcdCount = 8;

struct ResTile {
    int tx, ty;
};

static ResTile res_tile(int nx) {
    if (nx <= 32) return {32, 32};
    if (nx <= 64) return {64, 16};
    return {128, 8};
}

template <int TX, int TY, bool WRITE, typename Tu>
__global__ __launch_bounds__(TX* TY) __attribute__((amdgpu_waves_per_eu(8))) void residual_tile_kernel(
    const BatchPtr<const double> f, const BatchPtr<const Tu> u, float* __restrict__ r32,
    double* __restrict__ partial, const int nx, const int ny, const int nz, const double ix, const double iy,
    const double iz, const long bs, const int batch, const int chunk_gran, const int* __restrict__ done) {
    constexpr int NT = TX * TY, NW = NT / 64;
    __shared__ double t[2][TY + 2][TX + 2];
    const int gtx = (nx + TX - 1) / TX, gty = (ny + TY - 1) / TY;
    const int ngran = (nz + kZGran - 1) / kZGran, zc = (ngran + chunk_gran - 1) / chunk_gran;
    const int tiles = gtx * gty, per_problem = tiles * zc;
    const long total = (long)per_problem * batch;   // gridDim.x = per * kXcdCount >= total
    const long lin = blockIdx.x;
    const long per = (total + kXcdCount - 1) / kXcdCount;
    const long logical = (lin % kXcdCount) * per + lin / kXcdCount;
    if (logical >= total) return;
    const int p = (int)(logical / per_problem);
    const int blk = (int)(logical - (long)p * per_problem);   // chunk * tiles + tile
    if (done && done[p]) return;
    const int chunk = blk / tiles, tile = blk - chunk * tiles;
    const int x0 = (tile % gtx) * TX, y0 = (tile / gtx) * TY;
    const int z0 = chunk * chunk_gran * kZGran, z1 = min(nz, z0 + chunk_gran * kZGran);
    const long pl = (long)nx * ny;
    // buffer resources over problem p's fields: a load's address is a lane
    // byte offset (the column) plus a wave-uniform one (the plane), so the
    // per-plane part is one scalar multiply-add instead of 64-bit address
    // arithmetic per load
    const BufRange<double> fb(wave_uniform(f(p)), (unsigned)(pl * nz * (long)sizeof(double)));
    const BufRange<Tu> ub(wave_uniform(u(p)), (unsigned)(pl * nz * (long)sizeof(Tu)));
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const int x = x0 + tx, y = y0 + ty;
    const bool valid = x < nx && y < ny;
    // lanes outside the field read their nearest field point (lines their
    // neighbours read) and drop the value
    const long c0 = (long)(y < ny ? y : ny - 1) * nx + (x < nx ? x : nx - 1);
    // halo: the 2 TX + 2 TY cells around the tile (rows y0-1 and y0+TY, columns
    // x0-1 and x0+TX) go to threads 0 .. 2 TX + 2 TY - 1, one cell each; hOK
    // marks a cell inside the field (the others stay zero in LDS)
    static_assert(2 * TX + 2 * TY <= NT, "one halo cell per thread");
    const int hi = threadIdx.x;
    int hr = 0, hcol = 0, hx = -1, hy = -1;
    if (hi < TX) {
        hr = 0, hcol = hi + 1, hx = x0 + hi, hy = y0 - 1;
    } else if (hi < 2 * TX) {
        hr = TY + 1, hcol = hi - TX + 1, hx = x0 + hi - TX, hy = y0 + TY;
    } else if (hi < 2 * TX + TY) {
        hr = hi - 2 * TX + 1, hcol = 0, hx = x0 - 1, hy = y0 + hi - 2 * TX;
    } else if (hi < 2 * TX + 2 * TY) {
        hr = hi - 2 * TX - TY + 1, hcol = TX + 1, hx = x0 + TX, hy = y0 + hi - 2 * TX - TY;
    }
    const bool hOK = hx >= 0 && hx < nx && hy >= 0 && hy < ny;
    const int chy = hy < 0 ? 0 : (hy < ny ? hy : ny - 1), chx = hx < 0 ? 0 : (hx < nx ? hx : nx - 1);
    const long cH = hi < 2 * TX + 2 * TY ? (long)chy * nx + chx : c0;
    for (int i = threadIdx.x; i < 2 * (TY + 2) * (TX + 2); i += NT) (&t[0][0][0])[i] = 0.0;

    // Loads run D planes ahead and carry no branch: a plane index past the
    // field is clamped to its last plane (an in-bounds read), and the value is
    // replaced by the Dirichlet zero where it is used.  Values in flight stay
    // in their storage type Tu and are widened (and selected) at their use:
    // converting or selecting next to the load would make the wave wait there.
    // The in-flight values live in D + 1 register sets that trade roles from
    // plane to plane (the loop is unrolled by D + 1), so no register holding a
    // load in flight is ever copied: a copy would wait for the load.
    const int zl = nz - 1;
    const int plu = (int)(pl * (long)sizeof(Tu)), plf = (int)(pl * (long)sizeof(double));
    auto zcl = [&](int zz) { return zz < 0 ? 0 : (zz > zl ? zl : zz); };
    const int c0u = (int)(c0 * (long)sizeof(Tu)), c0f = (int)(c0 * (long)sizeof(double));
    const int cHu = (int)(cH * (long)sizeof(Tu));
    auto uval = [&](Tu v, int zz) { return (valid && zz >= 0 && zz < nz) ? (double)v : 0.0; };
    struct Slot {
        Tu u;       // u at z+1 (centre column)
        double f;   // f at z
        Tu h;       // the thread's halo cell at z
    };
    // FP64 iterates: only the waves that own halo cells load them (a
    // wave-uniform test; the norm-only residual 4% faster, the FP32-iterate
    // one measured no faster, so it keeps the unconditional load)
    const bool halo_wave =
        sizeof(Tu) != 8 || __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) * 64 < 2 * TX + 2 * TY;
    auto load_slot = [&](Slot& a, int zu, int zf) {
        a.u = ub.load(c0u, zcl(zu) * plu);
        a.f = fb.load(c0f, zcl(zf) * plf);
        if (halo_wave) a.h = ub.load(cHu, zcl(zf) * plu);
    };
    // prefetch depth in planes (three measured no faster than two for the
    // read-only FP64 pass: its memory parallelism is not what bounds it)
    constexpr int D = 2, NS = D + 1;
    const Tu um0 = ub.load(c0u, zcl(z0 - 1) * plu), uc0 = ub.load(c0u, zcl(z0) * plu);
    Slot sl[NS];
    static_for<0, D>([&](auto kk) {
        constexpr int q = decltype(kk)::value;
        load_slot(sl[q], z0 + q + 1, z0 + q);
    });
    double um = uval(um0, z0 - 1), uc = uval(uc0, z0);
    double rr = 0.0, ff = 0.0;
    __syncthreads();
    // one plane: `cur` holds u at z+1, f and halos at z; the next D-1 sets are
    // in flight; `fre` is refilled for D planes later
    auto step = [&](const int z, Slot& cur, Slot& fre) {
        const long o = (long)z * pl;
        auto& b = t[z & 1];
        b[ty + 1][tx + 1] = uc;
        if (hOK) b[hr][hcol] = (double)cur.h;
        load_slot(fre, z + D + 1, z + D);   // in flight while planes z .. z+D-1 are computed
        __syncthreads();   // b complete; the other buffer's readers (plane z-1) are past this point
        const double udn = uval(cur.u, z + 1);
        const double fv = cur.f;
        const double au = (b[ty + 1][tx] - 2.0 * uc + b[ty + 1][tx + 2]) * ix +
                          (b[ty][tx + 1] - 2.0 * uc + b[ty + 2][tx + 1]) * iy + (um - 2.0 * uc + udn) * iz;
        const double r = fv - au;
        if (valid) {
            rr += r * r;
            ff += fv * fv;
            if constexpr (WRITE) r32[p * bs + o + c0] = (float)r;
        }
        if ((z & (kZGran - 1)) == kZGran - 1 || z + 1 == nz) {   // granule done: the wave's partial
#pragma unroll
            for (int d = 32; d > 0; d >>= 1) {
                rr += __shfl_xor(rr, d, 64);
                ff += __shfl_xor(ff, d, 64);
            }
            if ((threadIdx.x & 63) == 0) {
                const long slot = ((long)p * ngran * tiles + (long)(z / kZGran) * tiles + tile) * NW + (threadIdx.x >> 6);
                partial[slot * 2 + 0] = rr;
                partial[slot * 2 + 1] = ff;
            }
            rr = 0.0;
            ff = 0.0;
        }
        um = uc;
        uc = udn;
    };
    int z = z0;
    for (; z + NS <= z1; z += NS)
        static_for<0, NS>([&](auto kk) {
            constexpr int q = decltype(kk)::value;
            step(z + q, sl[q], sl[(q + D) % NS]);
        });
    static_for<0, NS - 1>([&](auto kk) {   // the last z1 - z < NS planes
        constexpr int q = decltype(kk)::value;
        if (z + q < z1) step(z + q, sl[q], sl[(q + D) % NS]);
    });
}

// z-chunk length in granules: whole columns unless the batch is too small to
// give every CU a couple of blocks
static int res_chunk_gran(int nx, int ny, int nz, int batch, int tiles) {
    const int ngran = (nz + kZGran - 1) / kZGran;
    const long want = 2 * 256;
    const long have = (long)tiles * batch;
    if (have >= want) return ngran;
    const int zc = (int)std::min<long>(ngran, (want + have - 1) / have);
    return (ngran + zc - 1) / zc;
}

template <int TX, int TY, typename Tu>
static void launch_residual_tile(const ResidualArgs& a, const BatchPtr<const double>& f,
                                 const BatchPtr<const Tu>& u) {
    const int tiles = ((a.nx + TX - 1) / TX) * ((a.ny + TY - 1) / TY);
    const int cg = res_chunk_gran(a.nx, a.ny, a.nz, a.batch, tiles);
    const int ngran = (a.nz + kZGran - 1) / kZGran, zc = (ngran + cg - 1) / cg;
    // the kernel re-deals the blocks (problem, chunk, tile) to XCD slabs
    const long total = (long)tiles * zc * a.batch;
    const dim3 grid((unsigned)((total + kXcdCount - 1) / kXcdCount * kXcdCount));
    if (a.r32)
        hipLaunchKernelGGL((residual_tile_kernel<TX, TY, true, Tu>), grid, dim3(TX * TY), 0, a.stream, f, u, a.r32,
                           a.partial, a.nx, a.ny, a.nz, a.ix, a.iy, a.iz, a.bs, a.batch, cg, a.done);
    else
        hipLaunchKernelGGL((residual_tile_kernel<TX, TY, false, Tu>), grid, dim3(TX * TY), 0, a.stream, f, u, a.r32,
                           a.partial, a.nx, a.ny, a.nz, a.ix, a.iy, a.iz, a.bs, a.batch, cg, a.done);
}

// One block per problem: fixed-order sum of the partials, then the flag.  A
// thread adds partials tid, tid + kCheckBlock, ... with kCheckUnroll loads in
// flight (a 255^3 problem has 32768 partials), then the waves' sums are added
// in wave order; the order depends on nblk only, not on the batch.
constexpr int kCheckBlock = 1024;
constexpr int kCheckUnroll = 8;

// bad (the final certification of a solve, or null): set to 1 when a
// problem checked here ends above the tolerance (or its residual is NaN).
__global__ __launch_bounds__(kCheckBlock) void check_kernel(const double* __restrict__ partial, const int nblk,
                                                           const double tol, int* __restrict__ done,
                                                           int* __restrict__ iters, double* __restrict__ resid,
                                                           const int it, int* __restrict__ bad) {
    __shared__ double sh[2 * (kCheckBlock / 64)];
    const int p = blockIdx.x;
    if (done[p]) return;
    const auto part = as_global(partial + (long)p * nblk * 2);
    double rr = 0.0, ff = 0.0;
    for (int b = 0; b < nblk; b += kCheckBlock * kCheckUnroll) {
        double vr[kCheckUnroll], vf[kCheckUnroll];
#pragma unroll
        for (int k = 0; k < kCheckUnroll; ++k) {
            const int i = b + k * kCheckBlock + (int)threadIdx.x;
            const int ic = i < nblk ? i : nblk - 1;   // clamped: no branch around the loads
            const double x = part[2 * ic], y = part[2 * ic + 1];
            vr[k] = i < nblk ? x : 0.0;
            vf[k] = i < nblk ? y : 0.0;
        }
#pragma unroll
        for (int k = 0; k < kCheckUnroll; ++k) {
            rr += vr[k];
            ff += vf[k];
        }
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        rr += __shfl_xor(rr, d, 64);
        ff += __shfl_xor(ff, d, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sh[2 * w] = rr;
        sh[2 * w + 1] = ff;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, c = 0.0;
        for (int i = 0; i < kCheckBlock / 64; ++i) {
            a += sh[2 * i];
            c += sh[2 * i + 1];
        }
        const double rel = c > 0.0 ? sqrt(a / c) : sqrt(a);
        resid[p] = rel;
        iters[p] = it;
        if (rel <= tol) done[p] = 1;
        else if (bad) bad[0] = 1;
    }
}

__global__ void fill_int_kernel(int* __restrict__ a, int* __restrict__ b, const int n, const int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        a[i] = v;
        if (b) b[i] = v;
    }
}

constexpr long kResMaxBlocks = 1024;   // per problem

hipError_t launch_residual(const ResidualArgs& a) {
    const unsigned nblk = (unsigned)residual_partials(a.nx, a.ny, a.nz, a.nz > 1 || a.three);
    dim3 grid(nblk, a.batch);
    const BatchPtr<const double> f{a.f, a.f_tab, a.bs}, u{a.u, a.u_tab, a.bs};
    if (a.nz > 1 || a.three) {
        const ResTile t = res_tile(a.nx);
        if (a.u_f32) {
            const BatchPtr<const float> u32{(const float*)a.u, a.u_tab, a.bs};
            if (t.tx == 32) launch_residual_tile<32, 32>(a, f, u32);
            else if (t.tx == 64) launch_residual_tile<64, 16>(a, f, u32);
            else launch_residual_tile<128, 8>(a, f, u32);
        } else {
            if (t.tx == 32) launch_residual_tile<32, 32>(a, f, u);
            else if (t.tx == 64) launch_residual_tile<64, 16>(a, f, u);
            else launch_residual_tile<128, 8>(a, f, u);
        }
    } else if (a.u_f32) {
        return hipErrorInvalidValue;
    } else if (a.sy) {
        hipLaunchKernelGGL((residual_kernel<false, true>), grid, dim3(kResBlock), 0, a.stream, f, u, a.r32,
                           a.partial, a.nx, a.ny, 1, a.ix, a.iy, 0.0, a.sy, a.lx, a.ux, a.bs, a.done);
    } else {
        hipLaunchKernelGGL((residual_kernel<false, false>), grid, dim3(kResBlock), 0, a.stream, f, u, a.r32,
                           a.partial, a.nx, a.ny, 1, a.ix, a.iy, 0.0, nullptr, nullptr, nullptr, a.bs, a.done);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(check_kernel, dim3(a.batch), dim3(kCheckBlock), 0, a.stream, a.partial, (int)nblk, a.tol, a.done,
                       a.iters, a.resid, a.it, a.bad);
    return hipGetLastError();
}

hipError_t launch_check(const double* partial, int nblk, double tol, int* done, int* iters, double* resid, int it,
                        int batch, hipStream_t s) {
    hipLaunchKernelGGL(check_kernel, dim3(batch), dim3(kCheckBlock), 0, s, partial, nblk, tol, done, iters, resid, it,
                       (int*)nullptr);
    return hipGetLastError();
}

// partials per problem: 3-D, one per (z-granule, tile, wave) of
// residual_tile_kernel; 2-D, chunks of about 4 rows, a multiple of 8 (one
// slab per XCD), at most kResMaxBlocks
long residual_partials(int nx, int ny, int nz, bool three) {
    if (three) {
        const ResTile t = res_tile(nx);
        return (long)((nx + t.tx - 1) / t.tx) * ((ny + t.ty - 1) / t.ty) * ((nz + kZGran - 1) / kZGran) *
               (t.tx * t.ty / 64);
    }
    const long rows = (long)ny * nz;
    long b = (rows + 3) / 4;
    b = (b + 7) / 8 * 8;
    return b < kResMaxBlocks ? b : kResMaxBlocks;
}

// ---------------------------------------------------------------------------
// Periodic cubes (every axis periodic).  The operator is singular (it maps
// constants to zero) and the solve returns the zero-mean solution of
// A u = f - mean(f): the reference zeroes the (0,0,0) mode
// (src/lapl_cube.cpp:82-84).  Refinement therefore works on the projected
// residual r = (f - mean f) - A u, with A the 7-point stencil wrapping around
// every axis, and certifies ||r|| / ||f - mean f||.  mean f comes from a
// first pass; subtracting it before the squares keeps a large mean from
// swamping the small projected residual.  Each launch covers nblk contiguous
// chunks of a problem's points (block b: [b ch, (b+1) ch), ch = ceil(N/nblk))
// and reduces in a fixed order, so a result does not depend on the batch.
// ---------------------------------------------------------------------------
constexpr int kPerBlock = 256;

__device__ __forceinline__ void block_sum2_fixed(double& a, double& b, double* sh) {
    const int tid = threadIdx.x;
    sh[tid] = a;
    sh[kPerBlock + tid] = b;
    __syncthreads();
#pragma unroll
    for (int s = kPerBlock / 2; s > 0; s >>= 1) {
        if (tid < s) {
            sh[tid] += sh[tid + s];
            sh[kPerBlock + tid] += sh[kPerBlock + tid + s];
        }
        __syncthreads();
    }
    a = sh[0];
    b = sh[kPerBlock];
}

// partial[p][b][0] = sum of f over chunk b of problem p (in FP64)
template <typename T>
__global__ __launch_bounds__(kPerBlock) void fsum_kernel(const BatchPtr<const T> f, double* __restrict__ partial,
                                                         const long n, const int nblk, const int* __restrict__ done) {
    __shared__ double sh[2 * kPerBlock];
    const int p = blockIdx.y, b = blockIdx.x;
    if (done && done[p]) return;
    const long ch = (n + nblk - 1) / nblk, lo = b * ch, hi = lo + ch < n ? lo + ch : n;
    const auto fp = as_global(f(p));
    double s = 0.0, unused = 0.0;
    for (long i = lo + threadIdx.x; i < hi; i += kPerBlock) s += (double)fp[i];
    block_sum2_fixed(s, unused, sh);
    if (threadIdx.x == 0) {
        partial[((long)p * nblk + b) * 2] = s;
        partial[((long)p * nblk + b) * 2 + 1] = 0.0;
    }
}

// fmean[p] = (sum of the nblk chunk sums, in chunk order) / n
__global__ __launch_bounds__(kPerBlock) void fmean_kernel(const double* __restrict__ partial, const int nblk,
                                                          const long n, double* __restrict__ fmean,
                                                          const int* __restrict__ done) {
    __shared__ double sh[2 * kPerBlock];
    const int p = blockIdx.x;
    if (done && done[p]) return;
    double s = 0.0, unused = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kPerBlock) s += partial[((long)p * nblk + b) * 2];
    block_sum2_fixed(s, unused, sh);
    if (threadIdx.x == 0) fmean[p] = s / (double)n;
}

// r = (f - fmean) - A u (u == null: u = 0); r32 = (float) r when not null;
// partial[p][b] = (sum r^2, sum (f - fmean)^2) over chunk b
__global__ __launch_bounds__(kPerBlock) void residual_periodic_kernel(
    const BatchPtr<const double> f, const BatchPtr<const double> u, const double* __restrict__ fmean,
    float* __restrict__ r32, double* __restrict__ partial, const int nx, const int ny, const int nz,
    const double ix, const double iy, const double iz, const long bs, const int nblk, const int* __restrict__ done) {
    __shared__ double sh[2 * kPerBlock];
    const int p = blockIdx.y, b = blockIdx.x;
    if (done && done[p]) return;
    const long pl = (long)nx * ny, n = pl * nz;
    const long ch = (n + nblk - 1) / nblk, lo = b * ch, hi = lo + ch < n ? lo + ch : n;
    const auto fp = as_global(f(p));
    const double fm = fmean[p];
    const bool hu = u.base || u.tab;
    const double* up = hu ? u(p) : nullptr;
    double rr = 0.0, ff = 0.0;
    for (long c = lo + threadIdx.x; c < hi; c += kPerBlock) {
        const double fv = fp[c] - fm;
        double au = 0.0;
        if (hu) {
            const int x = (int)(c % nx), y = (int)((c / nx) % ny), z = (int)(c / pl);
            const long row = c - x, plane = c - (long)y * nx - x;
            const double uc = up[c];
            const double xm = up[row + (x == 0 ? nx - 1 : x - 1)], xp = up[row + (x == nx - 1 ? 0 : x + 1)];
            const double ym = up[plane + (long)(y == 0 ? ny - 1 : y - 1) * nx + x];
            const double yp = up[plane + (long)(y == ny - 1 ? 0 : y + 1) * nx + x];
            const double zm = up[(long)(z == 0 ? nz - 1 : z - 1) * pl + (c - (long)z * pl)];
            const double zp = up[(long)(z == nz - 1 ? 0 : z + 1) * pl + (c - (long)z * pl)];
            au = (xm - 2.0 * uc + xp) * ix + (ym - 2.0 * uc + yp) * iy + (zm - 2.0 * uc + zp) * iz;
        }
        const double r = fv - au;
        rr += r * r;
        ff += fv * fv;
        if (r32) r32[p * bs + c] = (float)r;
    }
    block_sum2_fixed(rr, ff, sh);
    if (threadIdx.x == 0) {
        partial[((long)p * nblk + b) * 2] = rr;
        partial[((long)p * nblk + b) * 2 + 1] = ff;
    }
}

// u = c - mean c (assign) or u += c - mean c, skipping converged problems.
// The FP32 inner solve zeroes the (0,0,0) mode only up to its rounding, and A
// cannot see a constant, so the refinement would keep the first solve's mean
// error; taking each correction's mean out keeps u zero-mean as the FP64
// solve returns it (and leaves every residual unchanged).
__global__ void update_f64_kernel(const BatchPtr<double> u, const float* __restrict__ c, const double* __restrict__ cmean,
                                  const long n, const long bs, const bool assign, const int* __restrict__ done) {
    const int p = blockIdx.y;
    if (done && done[p]) return;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const auto up = as_global(u(p));
    const double v = (double)c[p * bs + i] - cmean[p];
    up[i] = assign ? v : up[i] + v;
}

template <typename T>
static hipError_t launch_mean(const BatchPtr<const T>& f, double* out, const PeriodicResidualArgs& a) {
    const long n = (long)a.nx * a.ny * a.nz;
    hipLaunchKernelGGL(fsum_kernel<T>, dim3(a.nblk, a.batch), dim3(kPerBlock), 0, a.stream, f, a.partial, n, a.nblk,
                       a.done);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(fmean_kernel, dim3(a.batch), dim3(kPerBlock), 0, a.stream, (const double*)a.partial, a.nblk,
                       n, out, a.done);
    return hipGetLastError();
}

hipError_t launch_periodic_fmean(const PeriodicResidualArgs& a) {
    return launch_mean(BatchPtr<const double>{a.f, a.f_tab, a.bs}, a.fmean, a);
}

hipError_t launch_periodic_residual(const PeriodicResidualArgs& a) {
    const BatchPtr<const double> f{a.f, a.f_tab, a.bs}, u{a.u, a.u_tab, a.bs};
    hipLaunchKernelGGL(residual_periodic_kernel, dim3(a.nblk, a.batch), dim3(kPerBlock), 0, a.stream, f, u,
                       (const double*)a.fmean, a.r32, a.partial, a.nx, a.ny, a.nz, a.ix, a.iy, a.iz, a.bs, a.nblk,
                       a.done);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !a.check) return e;
    hipLaunchKernelGGL(check_kernel, dim3(a.batch), dim3(kCheckBlock), 0, a.stream, (const double*)a.partial, a.nblk,
                       a.tol, a.done, a.iters, a.resid, a.it, a.bad);
    return hipGetLastError();
}

hipError_t launch_periodic_update(const PeriodicResidualArgs& a, double* u, const void* const* u_tab, const float* c,
                                  bool assign) {
    hipError_t e = launch_mean(BatchPtr<const float>{c, nullptr, a.bs}, a.cmean, a);
    if (e != hipSuccess) return e;
    const long n = (long)a.nx * a.ny * a.nz;
    const dim3 g((unsigned)((n + 255) / 256), a.batch);
    hipLaunchKernelGGL(update_f64_kernel, g, dim3(256), 0, a.stream, BatchPtr<double>{u, u_tab, a.bs}, c,
                       (const double*)a.cmean, n, a.bs, assign, a.done);
    return hipGetLastError();
}

hipError_t launch_fill_int(int* a, int* b, int n, int v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_int_kernel, dim3((n + 255) / 256), dim3(256), 0, s, a, b, n, v);
    return hipGetLastError();
}

}  // namespace fdmgpu
