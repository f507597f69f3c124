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

// 3-D residual marching along z: a block of 64 x 4 threads covers
// x in [x0, x0+64), y in [y0, y0+4) and walks z; each thread keeps u at
// z-1, z, z+1 of its column in registers, so every u value is read from HBM
// once and the x / y neighbours come from L1/L2 (the neighbouring threads'
// loads of the same plane).  partial[p][blk] as residual_kernel, blk =
// blockIdx.y * gridDim.x + blockIdx.x.
constexpr int kZX = 64, kZY = 4;

template <bool WRITE>
__global__ __launch_bounds__(kZX* kZY) void residual_z_kernel(
    const BatchPtr<const double> f, const BatchPtr<const double> u, float* __restrict__ r32,
    double* __restrict__ partial, const int nx, const int ny, const int nz, const double ix, const double iy,
    const double iz, const long bs, const int* __restrict__ done) {
    __shared__ double sh[2 * kResBlock];
    const int p = blockIdx.z;
    if (done && done[p]) return;
    const auto fp = as_global(f(p));
    const auto up = as_global(u(p));
    const int x = blockIdx.x * kZX + (threadIdx.x & (kZX - 1));
    const int y = blockIdx.y * kZY + (threadIdx.x / kZX);
    const bool valid = x < nx && y < ny;
    const long pl = (long)nx * ny;
    const long c0 = (long)y * nx + x;
    const bool hx0 = x > 0, hx1 = x < nx - 1, hy0 = y > 0, hy1 = y < ny - 1;
    double rr = 0.0, ff = 0.0;
    if (valid) {
        double um = 0.0, uc = up[c0];
#pragma unroll 4
        for (int z = 0; z < nz; ++z) {
            const long c = z * pl + c0;
            const double un = z + 1 < nz ? up[c + pl] : 0.0;
            const double xm = hx0 ? up[c - 1] : 0.0, xp = hx1 ? up[c + 1] : 0.0;
            const double ym = hy0 ? up[c - nx] : 0.0, yp = hy1 ? up[c + nx] : 0.0;
            const double au = (xm - 2.0 * uc + xp) * ix + (ym - 2.0 * uc + yp) * iy + (um - 2.0 * uc + un) * iz;
            const double fv = fp[c];
            const double r = fv - au;
            if constexpr (WRITE) r32[p * bs + c] = (float)r;
            rr += r * r;
            ff += fv * fv;
            um = uc;
            uc = un;
        }
    }
    block_sum2<double>(rr, ff, sh);
    if (threadIdx.x == 0) {
        const int nblk = gridDim.x * gridDim.y;
        const int b = blockIdx.y * gridDim.x + blockIdx.x;
        partial[((long)p * nblk + b) * 2 + 0] = rr;
        partial[((long)p * nblk + b) * 2 + 1] = ff;
    }
}

// One block per problem: fixed-order sum of the partials, then the flag.
__global__ __launch_bounds__(kResBlock) void check_kernel(const double* __restrict__ partial, const int nblk,
                                                         const double tol, int* __restrict__ done,
                                                         int* __restrict__ iters, double* __restrict__ resid,
                                                         const int it) {
    __shared__ double sh[2 * kResBlock];
    const int p = blockIdx.x;
    if (done[p]) return;
    double rr = 0.0, ff = 0.0;
    for (int i = threadIdx.x; i < nblk; i += kResBlock) {
        rr += partial[((long)p * nblk + i) * 2 + 0];
        ff += partial[((long)p * nblk + i) * 2 + 1];
    }
    block_sum2<double>(rr, ff, sh);
    if (threadIdx.x == 0) {
        const double rel = ff > 0.0 ? sqrt(rr / ff) : sqrt(rr);
        resid[p] = rel;
        iters[p] = it;
        if (rel <= tol) done[p] = 1;
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
        dim3 zg((a.nx + kZX - 1) / kZX, (a.ny + kZY - 1) / kZY, a.batch);
        if (a.r32)
            hipLaunchKernelGGL((residual_z_kernel<true>), zg, dim3(kZX * kZY), 0, a.stream, f, u, a.r32, a.partial,
                               a.nx, a.ny, a.nz, a.ix, a.iy, a.iz, a.bs, a.done);
        else
            hipLaunchKernelGGL((residual_z_kernel<false>), zg, dim3(kZX * kZY), 0, a.stream, f, u, a.r32, a.partial,
                               a.nx, a.ny, a.nz, a.ix, a.iy, a.iz, a.bs, a.done);
    } else if (a.sy) {
        hipLaunchKernelGGL((residual_kernel<false, true>), grid, dim3(kResBlock), 0, a.stream, f, u, a.r32,
                           a.partial, a.nx, a.ny, 1, a.ix, a.iy, 0.0, a.sy, a.lx, a.ux, a.bs, a.done);
    } else {
        hipLaunchKernelGGL((residual_kernel<false, false>), grid, dim3(kResBlock), 0, a.stream, f, u, a.r32,
                           a.partial, a.nx, a.ny, 1, a.ix, a.iy, 0.0, nullptr, nullptr, nullptr, a.bs, a.done);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(check_kernel, dim3(a.batch), dim3(kResBlock), 0, a.stream, a.partial, (int)nblk, a.tol, a.done,
                       a.iters, a.resid, a.it);
    return hipGetLastError();
}

hipError_t launch_check(const double* partial, int nblk, double tol, int* done, int* iters, double* resid, int it,
                        int batch, hipStream_t s) {
    hipLaunchKernelGGL(check_kernel, dim3(batch), dim3(kResBlock), 0, s, partial, nblk, tol, done, iters, resid, it);
    return hipGetLastError();
}

// partials per problem: 3-D, one per residual_z_kernel block; 2-D, chunks of
// about 4 rows, a multiple of 8 (one slab per XCD), at most kResMaxBlocks
long residual_partials(int nx, int ny, int nz, bool three) {
    if (three) return (long)((nx + kZX - 1) / kZX) * ((ny + kZY - 1) / kZY);
    const long rows = (long)ny * nz;
    long b = (rows + 3) / 4;
    b = (b + 7) / 8 * 8;
    return b < kResMaxBlocks ? b : kResMaxBlocks;
}

hipError_t launch_fill_int(int* a, int* b, int n, int v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_int_kernel, dim3((n + 255) / 256), dim3(256), 0, s, a, b, n, v);
    return hipGetLastError();
}

}  // namespace fdmgpu
