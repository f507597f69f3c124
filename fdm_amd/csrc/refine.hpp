// FP64 residual / norm / convergence kernels for FP32+refinement (refine.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace fdmgpu {

struct ResidualArgs {
    const double* f;
    const double* u;
    float* r32;          // out: (float)(f - A u)
    double* partial;     // [batch][residual_partials()][2]
    int nx, ny, nz;
    bool three;          // 7-point (cube) vs 5-point (rect)
    double ix, iy, iz;   // 1/h^2 per axis of the inverted operator
    const double* sy;    // rect scales per column (lm_y_scale, L_scale, U_scale) or null
    const double* lx;
    const double* ux;
    long bs;
    int batch;
    double tol;
    int* done;
    int* iters;
    double* resid;
    int it;
    hipStream_t stream;
    const void* const* f_tab;   // per-problem pointers (device) or null
    const void* const* u_tab;
    bool u_f32 = false;         // u holds FP32 values (3-D only)
    int* bad = nullptr;         // final certification: set to 1 when a problem ends above tol
};

hipError_t launch_residual(const ResidualArgs& a);
// norm check over `nblk` partials per problem (fixed order); see refine.hip
hipError_t launch_check(const double* partial, int nblk, double tol, int* done, int* iters, double* resid, int it,
                        int batch, hipStream_t s);
// partials per problem of launch_residual (three: the 7-point cube operator)
long residual_partials(int nx, int ny, int nz, bool three);
// a[i] = b[i] = v for i < n (b may be null); a kernel, so it can be captured in a graph
hipError_t launch_fill_int(int* a, int* b, int n, int v, hipStream_t s);

// Periodic cubes (refine.hip): projected residual r = (f - mean f) - A u with
// the wrap-around 7-point stencil, over nblk chunks per problem.
struct PeriodicResidualArgs {
    const double* f;
    const double* u;             // null: u = 0 (the first right-hand side f - mean f)
    const void* const* f_tab;    // per-problem pointers (device) or null
    const void* const* u_tab;
    double* fmean;               // [batch]: mean of f (written by launch_periodic_fmean)
    double* cmean;               // [batch]: mean of the correction (launch_periodic_update)
    float* r32;                  // out: (float) r, or null
    double* partial;             // [batch][nblk][2]
    int nx, ny, nz, nblk;
    double ix, iy, iz;
    long bs;
    int batch;
    bool check;                  // run the norm check after the residual
    double tol;
    int* done;
    int* iters;
    double* resid;
    int it;
    hipStream_t stream;
    int* bad = nullptr;          // final certification: set to 1 when a problem ends above tol
};
hipError_t launch_periodic_fmean(const PeriodicResidualArgs& a);
hipError_t launch_periodic_residual(const PeriodicResidualArgs& a);
// u = c - mean c (assign) or u += c - mean c for the problems not done
// (c: FP32, [batch][elems] back to back)
hipError_t launch_periodic_update(const PeriodicResidualArgs& a, double* u, const void* const* u_tab, const float* c,
                                  bool assign);

}  // namespace fdmgpu
