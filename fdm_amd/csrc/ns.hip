// NSCube time step on the GPU (C ABI in include/fdmgpu.h: fdmgpu_ns_*).
//
// The reference's lid-driven cavity (src/ns_cube.h:12-78,
// src/ns_cube.cpp:27-277): staggered MAC grid, explicit advection + diffusion
// (F, G, H), pressure Poisson projection through LaplCube, velocity update.
// Every array lives in HBM with the reference's index ranges (x fastest):
//   u  i in [0,nz+1], k in [0,ny+1], j in [-1,nx+1]
//   v  i in [0,nz+1], k in [-1,ny+1], j in [0,nx+1]
//   w  i in [-1,nz+1], k in [0,ny+1], j in [0,nx+1]
//   p  [0,n+1]^3;  x, RHS [1,n]^3;  F j in [0,nx];  G k in [0,ny];  H i in [0,nz]
// A step is 4 kernels plus the Poisson solve on one stream; from the third
// step on it is replayed from a hipGraph captured from one step.
// Arithmetic is FP64 except inside an FP32+refine Poisson solve.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#include "kernels.hpp"
#include "plan.hpp"

using namespace fdmgpu;

namespace {

struct Grid {
    int nx, ny, nz;
    double dx, dy, dz, U0, Re, dt;
    // constant divisors of the stencils as reciprocals (set on the host):
    // 1/(Re h^2) per axis, 1/h per axis, dt/h per axis, 1/dt, 1/h^2 per axis
    double cx, cy, cz, idx, idy, idz, tdx, tdy, tdz, idt, idx2, idy2, idz2;
    // element offsets of (i, k, j) in each array
    __device__ long u(int i, int k, int j) const { return ((long)i * (ny + 2) + k) * (nx + 3) + (j + 1); }
    __device__ long v(int i, int k, int j) const { return ((long)i * (ny + 3) + (k + 1)) * (nx + 2) + j; }
    __device__ long w(int i, int k, int j) const { return ((long)(i + 1) * (ny + 2) + k) * (nx + 2) + j; }
    __device__ long p(int i, int k, int j) const { return ((long)i * (ny + 2) + k) * (nx + 2) + j; }
    __device__ long x(int i, int k, int j) const { return ((long)(i - 1) * ny + (k - 1)) * nx + (j - 1); }
    __device__ long F(int i, int k, int j) const { return ((long)(i - 1) * ny + (k - 1)) * (nx + 1) + j; }
    __device__ long G(int i, int k, int j) const { return ((long)(i - 1) * (ny + 1) + k) * nx + (j - 1); }
    __device__ long H(int i, int k, int j) const { return ((long)i * ny + (k - 1)) * nx + (j - 1); }
};

struct Fields {
    double *u, *v, *w, *p, *x, *F, *G, *H, *RHS;
};

__device__ __forceinline__ double sq(double a) { return a * a; }

// init_bound (src/ns_cube.cpp:65-122) in one launch: the lid, the ghost
// layers of u, v, w and the pressure on the faces.  The reference runs them
// in sequence: lid u[nz+1][k][j] = 2 U0 - u[nz][k][j] (k in [0, ny+1],
// j in [-1, nz+1]: the x bound is nz+1 in the reference), then the ghost
// copies, then the pressure faces from the ghosts.  Here every location is
// written by one thread with its final value, computed from values no thread
// of the launch writes:
//   task 0 (a = j, b = k): the lid row for j in [0, nz+1] other than the
//     ghost columns j = -1, nx+1 (the ghost copies overwrite those after the
//     lid in the reference);
//   task 1 (a = k, b = i): u's ghost columns -- at i = nz+1 the copy of the
//     lid-updated row, 2 U0 - u[nz][k][.], where the lid reached that column
//     -- and the pressure's x faces (i in [1, nz], k in [1, ny]) with the
//     ghost values as computed, not re-read;
//   task 2 (a = j, b = i): v's ghost rows and the y faces;
//   task 3 (a = j, b = k): w's ghost planes and the z faces.
// The pressure faces read only interior pressure (indices 1..n), which no
// thread writes.  The arithmetic is the reference's, so the results equal the
// three sequential passes bitwise.
__global__ void ns_bound_kernel(Grid g, double* __restrict__ u, double* __restrict__ v, double* __restrict__ w,
                                double* __restrict__ p) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const auto U = as_global(u);
    const auto V = as_global(v);
    const auto W = as_global(w);
    const auto P = as_global(p);
    if (blockIdx.z == 0) {   // lid: j = a - 1 in [-1, nz+1], k = b in [0, ny+1]
        const int j = a - 1, k = b;
        if (j > nz + 1 || k > ny + 1 || j == -1 || j == nx + 1) return;
        U[g.u(nz + 1, k, j)] = 2 * g.U0 - U[g.u(nz, k, j)];
    } else if (blockIdx.z == 1) {   // k = a in [0, ny+1], i = b in [0, nz+1]
        const int k = a, i = b;
        if (k > ny + 1 || i > nz + 1) return;
        const bool lid = i == nz + 1;
        const double gm = lid ? 2 * g.U0 - U[g.u(nz, k, 1)] : U[g.u(i, k, 1)];
        const double gp = lid ? (nx - 1 <= nz + 1 ? 2 * g.U0 - U[g.u(nz, k, nx - 1)] : U[g.u(nz + 1, k, nx - 1)])
                              : U[g.u(i, k, nx - 1)];
        U[g.u(i, k, -1)] = gm;
        U[g.u(i, k, nx + 1)] = gp;
        if (i >= 1 && i <= nz && k >= 1 && k <= ny) {
            P[g.p(i, k, 0)] = P[g.p(i, k, 1)] - (U[g.u(i, k, 1)] - 2 * U[g.u(i, k, 0)] + gm) / g.Re / g.dx;
            P[g.p(i, k, nx + 1)] = P[g.p(i, k, nx)] - (gp - 2 * U[g.u(i, k, nx)] + U[g.u(i, k, nx - 1)]) / g.Re / g.dx;
        }
    } else if (blockIdx.z == 2) {   // j = a in [0, nx+1], i = b in [0, nz+1]
        const int j = a, i = b;
        if (j > nx + 1 || i > nz + 1) return;
        const double gm = V[g.v(i, 1, j)], gp = V[g.v(i, ny - 1, j)];
        V[g.v(i, -1, j)] = gm;
        V[g.v(i, ny + 1, j)] = gp;
        if (i >= 1 && i <= nz && j >= 1 && j <= nx) {
            P[g.p(i, 0, j)] = P[g.p(i, 1, j)] - (V[g.v(i, 1, j)] - 2 * V[g.v(i, 0, j)] + gm) / g.Re / g.dy;
            P[g.p(i, ny + 1, j)] = P[g.p(i, ny, j)] - (gp - 2 * V[g.v(i, ny, j)] + V[g.v(i, ny - 1, j)]) / g.Re / g.dy;
        }
    } else {   // j = a in [0, nx+1], k = b in [0, ny+1]
        const int j = a, k = b;
        if (j > nx + 1 || k > ny + 1) return;
        const double gm = W[g.w(1, k, j)], gp = W[g.w(nz - 1, k, j)];
        W[g.w(-1, k, j)] = gm;
        W[g.w(nz + 1, k, j)] = gp;
        if (k >= 1 && k <= ny && j >= 1 && j <= nx) {
            P[g.p(0, k, j)] = P[g.p(1, k, j)] - (W[g.w(1, k, j)] - 2 * W[g.w(0, k, j)] + gm) / g.Re / g.dz;
            P[g.p(nz + 1, k, j)] = P[g.p(nz, k, j)] - (gp - 2 * W[g.w(nz, k, j)] + W[g.w(nz - 1, k, j)]) / g.Re / g.dz;
        }
    }
}

// Tiles of the stencil kernels (64 x 4 threads over x, y; one z-plane each)
// dealt to the XCDs in contiguous runs.  Workgroups of a 1-D grid go to the 8
// XCDs round-robin by blockIdx.x; XCD x takes tiles [x per, (x+1) per) in
// (x tile, y tile, z) order, so a tile's y and z neighbours, which read the
// same halo rows and planes, run on the same L2 close in time (with the tiles
// of a plane spread over all XCDs every XCD fetched the halo rows of its
// tiles and the planes above and below once more).  The grid is padded to a
// multiple of 8 blocks; blocks past the last tile return.
constexpr int kXcds = 8;
constexpr int kTileX = 64, kTileY = 4;

struct TileGrid {
    int gx, gy, gz;
    int blocks() const { return (gx * gy * gz + kXcds - 1) / kXcds * kXcds; }
};

inline TileGrid tile_grid(int nx, int ny, int nz) {
    return TileGrid{(nx + kTileX - 1) / kTileX, (ny + kTileY - 1) / kTileY, nz};
}

// Row of the tile a thread works on: a wavefront is one row of 64 threads
// (blockDim.x == kTileX), so the row is wave-uniform; said so, every row and
// plane part of an address is scalar arithmetic and a lane adds only its x.
static_assert(kTileX == 64, "a tile row is one wavefront");
__device__ __forceinline__ int tile_row(const int by) {
    return __builtin_amdgcn_readfirstlane(by * kTileY + (int)threadIdx.y);
}

__device__ __forceinline__ bool xcd_tile(const TileGrid tg, int& bx, int& by, int& bz) {
    const int per = (int)gridDim.x / kXcds;
    const int T = ((int)blockIdx.x % kXcds) * per + (int)blockIdx.x / kXcds;
    if (T >= tg.gx * tg.gy * tg.gz) return false;
    bx = T % tg.gx;
    const int r = T / tg.gx;
    by = r % tg.gy;
    bz = r / tg.gy;
    return true;
}

// F, G, H over the union box i in [0,nz], k in [0,ny], j in [0,nx]
// (src/ns_cube.cpp:131-201)
__global__ __launch_bounds__(kTileX * kTileY) void ns_fgh_kernel(Grid g, Fields f, const TileGrid tg) {
    int bx, by, bz;
    if (!xcd_tile(tg, bx, by, bz)) return;
    const int j = bx * kTileX + (int)threadIdx.x;
    const int k = tile_row(by);
    const int i = bz;
    if (j > g.nx || k > g.ny) return;
    const auto U = as_global((const double*)f.u);
    const auto V = as_global((const double*)f.v);
    const auto W = as_global((const double*)f.w);
    // every value F, G and H use, loaded once before any store (F, G, H
    // could alias u, v, w for all the compiler knows: loads after a store
    // would be issued again).  Neighbours outside an array's index range
    // (k - 1 < 0 for u and w, i - 1 < 0 for u and v, j - 1 < 0 for v and w)
    // are clamped into it; their values are used only where the unclamped
    // index is in range (F: i, k >= 1; G: i, j >= 1; H: k, j >= 1).
    const int im = i > 0 ? i - 1 : 0, km = k > 0 ? k - 1 : 0, jm = j > 0 ? j - 1 : 0;
    const double u0 = U[g.u(i, k, j)], uxp = U[g.u(i, k, j + 1)], uxm = U[g.u(i, k, j - 1)];
    const double uyp = U[g.u(i, k + 1, j)], uym = U[g.u(i, km, j)];
    const double uzp = U[g.u(i + 1, k, j)], uzm = U[g.u(im, k, j)];
    const double uypxm = U[g.u(i, k + 1, j - 1)], uzpxm = U[g.u(i + 1, k, j - 1)];
    const double v0 = V[g.v(i, k, j)], vxp = V[g.v(i, k, j + 1)], vxm = V[g.v(i, k, jm)];
    const double vyp = V[g.v(i, k + 1, j)], vym = V[g.v(i, k - 1, j)];
    const double vzp = V[g.v(i + 1, k, j)], vzm = V[g.v(im, k, j)];
    const double vymxp = V[g.v(i, k - 1, j + 1)], vzpym = V[g.v(i + 1, k - 1, j)];
    const double w0 = W[g.w(i, k, j)], wxp = W[g.w(i, k, j + 1)], wxm = W[g.w(i, k, jm)];
    const double wyp = W[g.w(i, k + 1, j)], wym = W[g.w(i, km, j)];
    const double wzp = W[g.w(i + 1, k, j)], wzm = W[g.w(i - 1, k, j)];
    const double wzmxp = W[g.w(i - 1, k, j + 1)], wzmyp = W[g.w(i - 1, k + 1, j)];
    const double dt = g.dt;
    if (i >= 1 && k >= 1) {
        as_global(f.F)[g.F(i, k, j)] =
            u0 + dt * ((uxp - 2 * u0 + uxm) * g.cx + (uyp - 2 * u0 + uym) * g.cy + (uzp - 2 * u0 + uzm) * g.cz -
                       (sq(0.5 * (u0 + uxp)) - sq(0.5 * (uxm + u0))) * g.idx -
                       0.25 * ((u0 + uyp) * (vxp + v0) - (uym + u0) * (vymxp + vym)) * g.idy -
                       0.25 * ((u0 + uzp) * (wxp + w0) - (uzm + u0) * (wzmxp + wzm)) * g.idz);
    }
    if (i >= 1 && j >= 1) {
        as_global(f.G)[g.G(i, k, j)] =
            v0 + dt * ((vxp - 2 * v0 + vxm) * g.cx + (vyp - 2 * v0 + vym) * g.cy + (vzp - 2 * v0 + vzm) * g.cz -
                       (sq(0.5 * (v0 + vyp)) - sq(0.5 * (vym + v0))) * g.idy -
                       0.25 * ((u0 + uyp) * (vxp + v0) - (uxm + uypxm) * (v0 + vxm)) * g.idx -
                       0.25 * ((w0 + wyp) * (v0 + vzp) - (wzm + wzmyp) * (vzm + v0)) * g.idz);
    }
    if (k >= 1 && j >= 1) {
        as_global(f.H)[g.H(i, k, j)] =
            w0 + dt * ((wxp - 2 * w0 + wxm) * g.cx + (wyp - 2 * w0 + wym) * g.cy + (wzp - 2 * w0 + wzm) * g.cz -
                       (sq(0.5 * (wzp + w0)) - sq(0.5 * (wzm + w0))) * g.idz -
                       0.25 * ((uzp + u0) * (wxp + w0) - (uzpxm + uxm) * (w0 + wxm)) * g.idx -
                       0.25 * ((w0 + wyp) * (v0 + vzp) - (wym + w0) * (vym + vzpym)) * g.idy);
    }
}

// Poisson right-hand side with the pressure boundary values moved over
// (src/ns_cube.cpp:204-235); interior i, k, j in [1, n]
__global__ __launch_bounds__(kTileX * kTileY) void ns_rhs_kernel(Grid g, Fields f, const TileGrid tg) {
    int bx, by, bz;
    if (!xcd_tile(tg, bx, by, bz)) return;
    const int j = bx * kTileX + (int)threadIdx.x + 1;
    const int k = tile_row(by) + 1;
    const int i = bz + 1;
    if (j > g.nx || k > g.ny) return;
    const auto Fa = as_global((const double*)f.F);
    const auto Ga = as_global((const double*)f.G);
    const auto Ha = as_global((const double*)f.H);
    const auto P = as_global((const double*)f.p);
    double r = ((Fa[g.F(i, k, j)] - Fa[g.F(i, k, j - 1)]) * g.idx + (Ga[g.G(i, k, j)] - Ga[g.G(i, k - 1, j)]) * g.idy +
                (Ha[g.H(i, k, j)] - Ha[g.H(i - 1, k, j)]) * g.idz) *
               g.idt;
    // the pressure's boundary values, read only at the points next to a face
    // (p has a ghost layer on every side, so every neighbour is in the array)
    if (i <= 1) r -= P[g.p(i - 1, k, j)] * g.idz2;
    if (k <= 1) r -= P[g.p(i, k - 1, j)] * g.idy2;
    if (j <= 1) r -= P[g.p(i, k, j - 1)] * g.idx2;
    if (j >= g.nx) r -= P[g.p(i, k, j + 1)] * g.idx2;
    if (k >= g.ny) r -= P[g.p(i, k + 1, j)] * g.idy2;
    if (i >= g.nz) r -= P[g.p(i + 1, k, j)] * g.idz2;
    as_global(f.RHS)[g.x(i, k, j)] = r;
}

// update_uvwp (src/ns_cube.cpp:238-277): interior faces from F, G, H and the
// pressure x; p's interior = x (p = x copies the index intersection)
__global__ __launch_bounds__(kTileX * kTileY) void ns_update_kernel(Grid g, Fields f, const TileGrid tg) {
    int bx, by, bz;
    if (!xcd_tile(tg, bx, by, bz)) return;
    const int j = bx * kTileX + (int)threadIdx.x + 1;
    const int k = tile_row(by) + 1;
    const int i = bz + 1;
    if (j > g.nx || k > g.ny) return;
    const auto X = as_global((const double*)f.x);
    // every load is issued before the first use (no branch around a load):
    // the x neighbours are clamped to the last interior point (their value
    // is used only below the bound); F, G and H hold the face at index n in
    // their own direction, so F(i,k,nx), G(i,ny,j), H(nz,k,j) are in the array
    const double xc = X[g.x(i, k, j)];
    const double xj = X[g.x(i, k, j < g.nx ? j + 1 : j)], xk = X[g.x(i, k < g.ny ? k + 1 : k, j)];
    const double xi = X[g.x(i < g.nz ? i + 1 : i, k, j)];
    const double fu = f.F[g.F(i, k, j)], fv = f.G[g.G(i, k, j)], fw = f.H[g.H(i, k, j)];
    if (j < g.nx) as_global(f.u)[g.u(i, k, j)] = fu - g.tdx * (xj - xc);
    if (k < g.ny) as_global(f.v)[g.v(i, k, j)] = fv - g.tdy * (xk - xc);
    if (i < g.nz) as_global(f.w)[g.w(i, k, j)] = fw - g.tdz * (xi - xc);
    as_global(f.p)[g.p(i, k, j)] = xc;
}


// Largest |value| of each of the nine arrays, ghost layers included, as
// tensor::maxabs (src/tensor.h:269-273): the running maximum starts at 0 and
// takes |b| only when m < |b|, so a NaN never enters it.  Pass 1: block
// (b, which) writes the maximum of its grid-stride share of array `which` to
// part[which][b]; pass 2: thread `which` folds that row.
constexpr int kMaxabsBlocks = 128, kMaxabsThreads = 256;

struct FieldSet {
    const double* a[9];
    unsigned long long n[9];
};

__device__ __forceinline__ double max_keep(double m, double b) { return m < b ? b : m; }

__global__ __launch_bounds__(kMaxabsThreads) void ns_maxabs_part_kernel(FieldSet fs, double* __restrict__ part) {
    __shared__ double wmax[kMaxabsThreads / 64];
    const int which = blockIdx.y;
    const auto a = as_global(fs.a[which]);
    const unsigned long long n = fs.n[which];
    double m = 0.0;
    for (unsigned long long i = (unsigned long long)blockIdx.x * kMaxabsThreads + threadIdx.x; i < n;
         i += (unsigned long long)kMaxabsBlocks * kMaxabsThreads)
        m = max_keep(m, fabs(a[i]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max_keep(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = wmax[0];
        for (int w = 1; w < kMaxabsThreads / 64; w++) r = max_keep(r, wmax[w]);
        part[which * kMaxabsBlocks + blockIdx.x] = r;
    }
}

__global__ void ns_maxabs_final_kernel(const double* __restrict__ part, double* __restrict__ out) {
    const int which = threadIdx.x;
    if (which >= 9) return;
    double m = 0.0;
    for (int b = 0; b < kMaxabsBlocks; b++) m = max_keep(m, part[which * kMaxabsBlocks + b]);
    out[which] = m;
}

}  // namespace

struct fdmgpu_ns_t {
    int device = 0;
    Grid g{};
    Fields f{};
    size_t count[9] = {};
    fdmgpu_plan plan = nullptr;
    hipStream_t cap = nullptr;     // capture stream
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    long steps = 0;
    bool warm_ok = false;          // x holds the pressure of the previous step (a field upload clears it)
    bool use_graph = true;
    int* d_bad = nullptr;          // set on the device when a projection misses its tolerance (sticky)
    double* d_max = nullptr;       // maxabs: 9 x kMaxabsBlocks partial maxima, then the 9 results
    hipEvent_t ev_phase[5] = {};   // timed steps: the bounds of the four phases
};

namespace {

#define NS_HIPCHK(x)                                                                     \
    do {                                                                                 \
        hipError_t _e = (x);                                                             \
        if (_e != hipSuccess) {                                                          \
            set_error_text(hipGetErrorString(_e));                                       \
            return _e == hipErrorOutOfMemory ? FDMGPU_ERR_OUT_OF_MEMORY : FDMGPU_ERR_DEVICE; \
        }                                                                                \
    } while (0)

// One step on stream s.  ev (a timed step, never under capture): events
// recorded at the bounds of the reference's four phases, init_bound (lid,
// ghost layers, pressure boundary), FGH, poisson (right-hand side and solve)
// and update_uvwp (src/ns_cube.cpp:27-50).
int ns_enqueue_step(fdmgpu_ns_t* N, hipStream_t s, hipEvent_t* ev = nullptr) {
    const Grid& g = N->g;
    const dim3 b2(kTileX, kTileY);
    auto mark = [&](int i) { return ev ? hipEventRecord(ev[i], s) : hipSuccess; };
    (void)hipGetLastError();
    NS_HIPCHK(mark(0));
    {
        const int fa = std::max(std::max(g.nz + 3, g.ny + 2), g.nx + 2), fb = std::max(g.nz, g.ny) + 2;
        hipLaunchKernelGGL(ns_bound_kernel, dim3((fa + 63) / 64, fb, 4), dim3(64), 0, s, g, N->f.u, N->f.v, N->f.w,
                           N->f.p);
        NS_HIPCHK(hipGetLastError());
    }
    NS_HIPCHK(mark(1));
    const TileGrid tf = tile_grid(g.nx + 1, g.ny + 1, g.nz + 1), ti = tile_grid(g.nx, g.ny, g.nz);
    hipLaunchKernelGGL(ns_fgh_kernel, dim3(tf.blocks()), b2, 0, s, g, N->f, tf);
    NS_HIPCHK(hipGetLastError());
    NS_HIPCHK(mark(2));
    hipLaunchKernelGGL(ns_rhs_kernel, dim3(ti.blocks()), b2, 0, s, g, N->f, ti);
    NS_HIPCHK(hipGetLastError());
    // FP32+refine: from the second step on, x still holds the previous step's
    // pressure, a close iterate for the new right-hand side; the refinement
    // continues from it (FP64 residual, then FP32 corrections: inner solves
    // 2 and, as a guard that costs only empty launches once converged, 3)
    // instead of solving from zero.  The tolerance and the stopping test are
    // those of a cold solve.
    const bool refine = N->plan->precision == FDMGPU_F32_REFINE64;
    const bool warm = refine && N->warm_ok;
    // a projection that misses the tolerance marks the run: the solve's final
    // norm check sets d_bad on the device (no host round trip, no extra
    // launch); only this step's solves carry the flag
    N->plan->w.bad = N->d_bad;
    int rc = warm ? plan_solve_iters(N->plan, N->plan->w, N->f.x, nullptr, N->f.RHS, nullptr, 1, s, 2,
                                     std::max(2, N->plan->max_iter - 1))
                  : plan_solve(N->plan, N->plan->w, N->f.x, nullptr, N->f.RHS, nullptr, 1, s);
    N->plan->w.bad = nullptr;
    if (rc) return rc;
    NS_HIPCHK(mark(3));
    hipLaunchKernelGGL(ns_update_kernel, dim3(ti.blocks()), b2, 0, s, g, N->f, ti);
    NS_HIPCHK(hipGetLastError());
    NS_HIPCHK(mark(4));
    return FDMGPU_OK;
}

// fdmgpu_get_stats on the solver plan reports the last step's pressure solve
void ns_note_steps(fdmgpu_ns_t* N, hipStream_t s) {
    if (N->steps == 0) return;
    N->plan->last_stream = s;
    N->plan->last_count = 1;
    N->plan->last_iters = N->plan->w.iters;
    N->plan->last_resid = N->plan->w.resid;
}

void ns_free(fdmgpu_ns_t* N) {
    if (!N) return;
    (void)hipSetDevice(N->device);
    if (N->cap) (void)hipStreamSynchronize(N->cap);
    if (N->exec) (void)hipGraphExecDestroy(N->exec);
    if (N->graph) (void)hipGraphDestroy(N->graph);
    double* a[] = {N->f.u, N->f.v, N->f.w, N->f.p, N->f.x, N->f.F, N->f.G, N->f.H, N->f.RHS};
    for (double* p : a) (void)hipFree(p);
    (void)hipFree(N->d_bad);
    (void)hipFree(N->d_max);
    for (hipEvent_t e : N->ev_phase)
        if (e) (void)hipEventDestroy(e);
    if (N->plan) fdmgpu_destroy(N->plan);
    if (N->ev_in) (void)hipEventDestroy(N->ev_in);
    if (N->ev_out) (void)hipEventDestroy(N->ev_out);
    if (N->cap) (void)hipStreamDestroy(N->cap);
    delete N;
}

double** ns_field(fdmgpu_ns_t* N, int which) {
    double** a[] = {&N->f.u, &N->f.v, &N->f.w, &N->f.p, &N->f.x, &N->f.F, &N->f.G, &N->f.H, &N->f.RHS};
    return (which >= 0 && which < 9) ? a[which] : nullptr;
}

}  // namespace

extern "C" {

int fdmgpu_ns_create(fdmgpu_ns* ns, int solver_precision, int nx, int ny, int nz, double x1, double x2, double y1,
                     double y2, double z1, double z2, double u0, double Re, double dt) {
    if (!ns || (solver_precision != FDMGPU_F64 && solver_precision != FDMGPU_F32_REFINE64)) return FDMGPU_ERR_INVALID_ARG;
    *ns = nullptr;
    if (nx < 2 || ny < 2 || nz < 2) return FDMGPU_ERR_INVALID_SIZE;
    if (nz > nx) return FDMGPU_ERR_UNSUPPORTED;   // the lid loop's x bound is nz+1 (src/ns_cube.cpp:68)
    if (!(Re > 0) || !(dt > 0) || !(x2 > x1) || !(y2 > y1) || !(z2 > z1)) return FDMGPU_ERR_INVALID_ARG;
    fdmgpu_ns_t* N = new fdmgpu_ns_t();
    if (hipGetDevice(&N->device) != hipSuccess) {
        delete N;
        return FDMGPU_ERR_DEVICE;
    }
    Grid& g = N->g;
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.dx = (x2 - x1) / nx; g.dy = (y2 - y1) / ny; g.dz = (z2 - z1) / nz;
    g.U0 = u0; g.Re = Re; g.dt = dt;
    g.cx = 1.0 / (Re * g.dx * g.dx); g.cy = 1.0 / (Re * g.dy * g.dy); g.cz = 1.0 / (Re * g.dz * g.dz);
    g.idx = 1.0 / g.dx; g.idy = 1.0 / g.dy; g.idz = 1.0 / g.dz;
    g.tdx = dt / g.dx; g.tdy = dt / g.dy; g.tdz = dt / g.dz;
    g.idt = 1.0 / dt;
    g.idx2 = 1.0 / (g.dx * g.dx); g.idy2 = 1.0 / (g.dy * g.dy); g.idz2 = 1.0 / (g.dz * g.dz);
    const size_t X = nx, Y = ny, Z = nz;
    const size_t cnt[9] = {(Z + 2) * (Y + 2) * (X + 3), (Z + 2) * (Y + 3) * (X + 2), (Z + 3) * (Y + 2) * (X + 2),
                           (Z + 2) * (Y + 2) * (X + 2), Z * Y * X, Z * Y * (X + 1), Z * (Y + 1) * X,
                           (Z + 1) * Y * X, Z * Y * X};
    int rc = FDMGPU_OK;
    for (int i = 0; i < 9 && !rc; i++) {
        N->count[i] = cnt[i];
        double** p = ns_field(N, i);
        if (hipMalloc(p, sizeof(double) * cnt[i]) != hipSuccess ||
            hipMemset(*p, 0, sizeof(double) * cnt[i]) != hipSuccess)   // fluid at rest (zero-initialised tensors)
            rc = FDMGPU_ERR_OUT_OF_MEMORY;
    }
    // the solver of NSCube::lapl_solver (src/ns_cube.h:77)
    if (!rc) rc = fdmgpu_plan_cube(&N->plan, solver_precision, 0, g.dx, g.dy, g.dz, x2 - x1 + g.dx, y2 - y1 + g.dy,
                                   z2 - z1 + g.dz, nx, ny, nz);
    if (!rc && (hipMalloc(&N->d_bad, sizeof(int)) != hipSuccess || hipMemset(N->d_bad, 0, sizeof(int)) != hipSuccess))
        rc = FDMGPU_ERR_OUT_OF_MEMORY;
    if (!rc && (hipStreamCreateWithFlags(&N->cap, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&N->ev_in, hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&N->ev_out, hipEventDisableTiming) != hipSuccess))
        rc = FDMGPU_ERR_DEVICE;
    if (!rc && hipDeviceSynchronize() != hipSuccess) rc = FDMGPU_ERR_DEVICE;
    if (rc) {
        ns_free(N);
        return rc;
    }
    *ns = N;
    return FDMGPU_OK;
}

int fdmgpu_ns_step(fdmgpu_ns N, int count, void* stream) {
    if (!N || count < 0) return FDMGPU_ERR_INVALID_ARG;
    NS_HIPCHK(hipSetDevice(N->device));
    hipStream_t user = (hipStream_t)stream;
    hipStream_t s = N->cap;
    NS_HIPCHK(hipEventRecord(N->ev_in, user));
    NS_HIPCHK(hipStreamWaitEvent(s, N->ev_in, 0));
    for (int c = 0; c < count; c++) {
        // eager for the first two steps (allocations, warm-up) and for a step
        // whose solve starts from zero after a field upload; replayed after
        if (!N->use_graph || N->steps < 2 || !N->warm_ok) {
            int rc = ns_enqueue_step(N, s);
            if (rc) return rc;
        } else {
            if (!N->exec) {
                NS_HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
                int rc = ns_enqueue_step(N, s);
                hipGraph_t graph = nullptr;
                hipError_t e = hipStreamEndCapture(s, &graph);
                if (!rc && e != hipSuccess) rc = FDMGPU_ERR_DEVICE;
                if (rc) {
                    if (graph) (void)hipGraphDestroy(graph);
                    return rc;
                }
                N->graph = graph;
                NS_HIPCHK(hipGraphInstantiate(&N->exec, N->graph, nullptr, nullptr, 0));
            }
            NS_HIPCHK(hipGraphLaunch(N->exec, s));
        }
        N->steps++;
        N->warm_ok = true;
    }
    ns_note_steps(N, s);
    NS_HIPCHK(hipEventRecord(N->ev_out, s));
    NS_HIPCHK(hipStreamWaitEvent(user, N->ev_out, 0));
    return FDMGPU_OK;
}

int fdmgpu_ns_step_timed(fdmgpu_ns N, double* ms, void* stream) {
    if (!N || !ms) return FDMGPU_ERR_INVALID_ARG;
    NS_HIPCHK(hipSetDevice(N->device));
    for (hipEvent_t& e : N->ev_phase)
        if (!e) NS_HIPCHK(hipEventCreate(&e));
    hipStream_t user = (hipStream_t)stream;
    hipStream_t s = N->cap;
    NS_HIPCHK(hipEventRecord(N->ev_in, user));
    NS_HIPCHK(hipStreamWaitEvent(s, N->ev_in, 0));
    // eager: the same kernels with the same arguments as a replayed step
    int rc = ns_enqueue_step(N, s, N->ev_phase);
    if (rc) return rc;
    N->steps++;
    N->warm_ok = true;
    ns_note_steps(N, s);
    NS_HIPCHK(hipEventRecord(N->ev_out, s));
    NS_HIPCHK(hipStreamWaitEvent(user, N->ev_out, 0));
    NS_HIPCHK(hipEventSynchronize(N->ev_phase[4]));
    for (int i = 0; i < 4; i++) {
        float t = 0.f;
        NS_HIPCHK(hipEventElapsedTime(&t, N->ev_phase[i], N->ev_phase[i + 1]));
        ms[i] = t;
    }
    return FDMGPU_OK;
}

int fdmgpu_ns_maxabs(fdmgpu_ns N, double* out) {
    if (!N || !out) return FDMGPU_ERR_INVALID_ARG;
    NS_HIPCHK(hipSetDevice(N->device));
    if (!N->d_max) NS_HIPCHK(hipMalloc(&N->d_max, sizeof(double) * 9 * (kMaxabsBlocks + 1)));
    FieldSet fs{};
    for (int i = 0; i < 9; i++) {
        fs.a[i] = *ns_field(N, i);
        fs.n[i] = N->count[i];
    }
    hipStream_t s = N->cap;
    double* res = N->d_max + 9 * kMaxabsBlocks;
    (void)hipGetLastError();
    hipLaunchKernelGGL(ns_maxabs_part_kernel, dim3(kMaxabsBlocks, 9), dim3(kMaxabsThreads), 0, s, fs, N->d_max);
    NS_HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(ns_maxabs_final_kernel, dim3(1), dim3(64), 0, s, (const double*)N->d_max, res);
    NS_HIPCHK(hipGetLastError());
    NS_HIPCHK(hipMemcpyAsync(out, res, sizeof(double) * 9, hipMemcpyDeviceToHost, s));
    NS_HIPCHK(hipStreamSynchronize(s));
    return FDMGPU_OK;
}

int fdmgpu_ns_get(fdmgpu_ns N, int which, double* host, unsigned long long count) {
    if (!N || !host) return FDMGPU_ERR_INVALID_ARG;
    double** p = ns_field(N, which);
    if (!p || count != N->count[which]) return FDMGPU_ERR_INVALID_ARG;
    NS_HIPCHK(hipSetDevice(N->device));
    NS_HIPCHK(hipStreamSynchronize(N->cap));
    NS_HIPCHK(hipMemcpy(host, *p, sizeof(double) * count, hipMemcpyDeviceToHost));
    int bad = 0;
    NS_HIPCHK(hipMemcpy(&bad, N->d_bad, sizeof(int), hipMemcpyDeviceToHost));
    return bad ? FDMGPU_ERR_NOT_CONVERGED : FDMGPU_OK;
}

int fdmgpu_ns_set(fdmgpu_ns N, int which, const double* host, unsigned long long count) {
    if (!N || !host) return FDMGPU_ERR_INVALID_ARG;
    double** p = ns_field(N, which);
    if (!p || count != N->count[which]) return FDMGPU_ERR_INVALID_ARG;
    NS_HIPCHK(hipSetDevice(N->device));
    NS_HIPCHK(hipStreamSynchronize(N->cap));
    NS_HIPCHK(hipMemcpy(*p, host, sizeof(double) * count, hipMemcpyHostToDevice));
    N->warm_ok = false;   // the next pressure solve starts from zero
    return FDMGPU_OK;
}

int fdmgpu_ns_info(fdmgpu_ns N, int which, unsigned long long* count, void** device_ptr) {
    if (!N) return FDMGPU_ERR_INVALID_ARG;
    double** p = ns_field(N, which);
    if (!p) return FDMGPU_ERR_INVALID_ARG;
    if (count) *count = N->count[which];
    if (device_ptr) *device_ptr = *p;
    return FDMGPU_OK;
}

int fdmgpu_ns_solver(fdmgpu_ns N, fdmgpu_plan* plan) {
    if (!N || !plan) return FDMGPU_ERR_INVALID_ARG;
    *plan = N->plan;
    return FDMGPU_OK;
}

int fdmgpu_ns_destroy(fdmgpu_ns N) {
    if (!N) return FDMGPU_ERR_INVALID_ARG;
    ns_free(N);
    return FDMGPU_OK;
}

}  // extern "C"
