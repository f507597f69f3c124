// Internal interface between the C ABI (fdmgpu.hip) and the batching /
// serving layer (serve.hip): solver plans and per-solve scratch.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "../../include/fdmgpu.h"

namespace fdmgpu {

// Scratch of one in-flight solve of up to `cap` problems (FP32+refine plans):
// the FP32 correction field, norm partials and the per-problem convergence
// state.  A plan owns one for its own entry points; the batching layer gives
// every in-flight batch its own, so batches of one plan can overlap.
struct Work {
    float* r32 = nullptr;               // [cap][elems]
    float* u32 = nullptr;               // [cap][elems] cubes: the iterate after inner solve 1 (FP32 values)
    double* partial = nullptr;          // [cap][nblk][2]
    int* done = nullptr;                // [cap]
    int* iters = nullptr;               // [cap]
    double* resid = nullptr;            // [cap]
    int cap = 0;
    double* fcopy = nullptr;            // f kept when a refine solve runs in place
    size_t fcopy_bytes = 0;
    int* bad = nullptr;                 // device flag set when a solve ends above tolerance (or null)
};

using Tab = const void* const*;

// Kernel variant of a plan (chosen by the autotuner; 0 = built-in default).
struct Variant {
    int plane = 1;      // square planes that fit LDS: 1 = plane kernel, 0 = x and y tile kernels,
                        // 2 = streaming plane kernel (planes of N = 128: FP64, the default
                        // there; FP32 and the refinement's FP32-to-FP32 transforms)
    int tile_x = 0;     // x tile kernel: rows per block (16/32/64)
    int tile_y = 0;     // y tile kernel: columns per block
    int zcr_gmul = 0;   // z cyclic reduction: lanes-per-system multiplier (1/2)
    int fuse = 0;       // FP32+refine on the plane path: 1 = residual fused into the forward
                        // plane transform, 0 = separate residual pass writing the FP32 rhs
                        // (the default: the tiled residual kernel streams faster than the
                        // plane kernel computes the stencil)
};

}  // namespace fdmgpu

struct fdmgpu_plan_t {
    int kind = 3;          // 3 = LaplCube, 2 = LaplRect
    int precision = FDMGPU_F64;
    int periodic = 0;
    int kfwd = 0, kinv = 0;             // pencil transform kinds (kDst, or kRdftFwd/kRdftInv if periodic)
    bool fft2 = false;                  // LaplRectFFT2: x transformed too (kinds kxf/kxi), spectral divide
    int kxf = 0, kxi = 0;
    int nx = 0, ny = 0, nz = 1;
    int logx = 0, logy = 0, logz = 0;   // log2 of the transform lengths
    double dx = 1, dy = 1, dz = 1, lx = 1, ly = 1, lz = 1;
    double hx_op = 1, hz_op = 1;        // spacings of the inverted operator (eigenvalue aliasing)
    int device = 0;
    size_t elem = 8;                    // bytes per element at the API
    size_t csize = 8;                   // bytes per compute element
    void* tab_x = nullptr;              // DST tables (compute type), x axis
    void* tab_y = nullptr;              // y axis (== tab_x when the lengths match)
    void* lmx = nullptr;                // eigenvalues per mode (compute type)
    void* lmy = nullptr;
    // LaplRect scales, per column (index 0..nx-1 = reference index 1..nx)
    std::vector<double> h_sc, h_Ls, h_Us;
    std::vector<double> h_lmx;          // LaplRectFFT2: x eigenvalues per mode
    void* d_sc = nullptr;               // compute type
    void* d_Ls = nullptr;
    void* d_Us = nullptr;
    double* d_sc64 = nullptr;           // FP64 copies for the residual
    double* d_Ls64 = nullptr;
    double* d_Us64 = nullptr;
    bool scaled = false;
    bool pivot_rows = false;            // LaplRect rows not all diagonally dominant: ?gtsv with pivoting
    // refinement
    double rtol = 1e-10;
    int max_iter = 4;
    fdmgpu::Work w;                     // scratch of the plan's own solves
    fdmgpu::Variant var;                // kernel variant
    // host staging
    void* stage = nullptr;
    size_t stage_bytes = 0;
    hipStream_t own = nullptr;
    hipStream_t last_stream = nullptr;
    int last_count = 0;
    const int* last_iters = nullptr;    // per-problem results of the last refine solve (device):
    const double* last_resid = nullptr; //   in the scratch that solve used
    // per-kernel timing: bit k set = launches of kernel class k are timed
    // (fdmgpu_set_timing, fdmgpu_set_timing_kernels)
    unsigned timing = 0;
    std::vector<hipEvent_t> evs;        // 2 per recorded launch
    std::vector<int> rec_kernel;
    size_t nrec = 0;
    double tot_ms[FDMGPU_NUM_KERNELS] = {};
    int launches[FDMGPU_NUM_KERNELS] = {};
};

namespace fdmgpu {

// Device bytes of the scratch a solve of `count` problems needs (each array
// 256-byte aligned), for a plan or for a shape and precision.
size_t work_bytes(const fdmgpu_plan_t* P, int count);
size_t work_bytes_shape(int kind, int precision, int nx, int ny, int nz, int count);
// A Work over caller-provided memory of work_bytes(P, count) bytes (never freed
// by work_free's owner: do not pass it to work_free).
Work work_carve(const fdmgpu_plan_t* P, int count, void* mem);
// Grows W to `count` problems (no-op when large enough).
int work_reserve(const fdmgpu_plan_t* P, Work& W, int count);
void work_free(Work& W);
// Enqueues the solve of `count` problems on stream s: problem p's arrays at
// ans + p*elems / rhs + p*elems, or at ans_tab[p] / rhs_tab[p] (device tables).
int plan_solve(fdmgpu_plan_t* P, Work& W, void* ans, Tab ans_tab, const void* rhs, Tab rhs_tab, int count,
               hipStream_t s);
// FP32+refine plans: inner solves first..last only (first > 1 continues from
// the iterate in ans, whose residual after first-1 solves missed the tolerance).
int plan_solve_iters(fdmgpu_plan_t* P, Work& W, void* ans, Tab ans_tab, const void* rhs, Tab rhs_tab, int count,
                     hipStream_t s, int first, int last);
// Elements of one problem.
inline long plan_elems(const fdmgpu_plan_t* P) { return (long)P->nx * P->ny * P->nz; }
// Error text of the last failure on this thread (FDMGPU_ERR_DEVICE details).
void set_error_text(const char* msg);
// Whether plan P can use the plane kernel (square planes that fit LDS).
bool plane_capable(const fdmgpu_plan_t* P);
// plane field of P's default variant (0 axis tiles, 1 plane kernel, 2 streaming plane kernel)
int default_plane(const fdmgpu_plan_t* P);
// Kernel variants valid for P (the first is the default).
std::vector<Variant> plan_variants(const fdmgpu_plan_t* P);
bool variant_valid(const fdmgpu_plan_t* P, const Variant& v);
// Applies the autotuner's choice for P's shape, if one is known (serve.hip).
void apply_tuned(fdmgpu_plan_t* P);

}  // namespace fdmgpu
