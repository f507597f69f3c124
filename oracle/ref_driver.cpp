// Thin C entry points over the REFERENCE's own code, for oracle/_ref only.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the reference's sources where they lie (src/fft.cpp, src/lapl_cube.cpp,
// src/asp_gauss.cpp, src/asp_fft.cpp, src/asp_misc.cpp under the reference
// checkout); nothing here is copied from the reference.  The result,
// oracle/_ref/libfdmref.so, is used to pin oracle/fdm_oracle.cpp and as the
// "reference" CPU baseline in bench.py.  src/lapl_rect.cpp is not built: it
// needs a LAPACK providing dgtsv_/sgtsv_, which this image does not have.

#include <cstring>
#include <vector>

#include "asp_gauss.h"
#include "cyclic_reduction.h"
#include "fft.h"
#include "lapl_cube.h"

void sFT(double* S, double* s, double dx, int N);   // src/asp_fft.cpp:308
void cFT(double* S, double* s, double dx, int N);   // src/asp_fft.cpp:404

namespace {
template <typename T>
void sfft(T* S, const T* s, int N, T dx, int kind) {
    fdm::FFTTable<T> table(N);
    fdm::FFT<T> ft(table, N);
    std::vector<T> in(s, s + N + 1);   // the transforms clobber their input
    if (kind == 0) ft.sFFT(S, in.data(), dx);
    else if (kind == 1) ft.cFFT(S, in.data(), dx);
    else if (kind == 2) ft.pFFT_1(S, in.data(), dx);
    else ft.pFFT(S, in.data(), dx);
}

using per3 = fdm::tensor_flags<fdm::tensor_flag::periodic, fdm::tensor_flag::periodic,
                               fdm::tensor_flag::periodic>;

template <typename T>
void cube(T* ans, const T* rhs, int periodic, double dx, double dy, double dz, double lx, double ly,
          double lz, int nx, int ny, int nz) {
    std::vector<T> r(rhs, rhs + (size_t)nx * ny * nz);
    if (periodic) {
        fdm::LaplCube<T, false, per3> l(dx, dy, dz, lx, ly, lz, nx, ny, nz);
        l.solve(ans, r.data());
    } else {
        fdm::LaplCube<T, false> l(dx, dy, dz, lx, ly, lz, nx, ny, nz);
        l.solve(ans, r.data());
    }
}

// A constructed solver kept alive for timing repeated solves.
struct CubeHandle {
    virtual ~CubeHandle() {}
    virtual void solve(double* ans, double* rhs) = 0;
};
struct CubeD : CubeHandle {
    fdm::LaplCube<double, false> l;
    CubeD(double dx, double dy, double dz, double lx, double ly, double lz, int nx, int ny, int nz)
        : l(dx, dy, dz, lx, ly, lz, nx, ny, nz) {}
    void solve(double* ans, double* rhs) override { l.solve(ans, rhs); }
};
}  // namespace

extern "C" {
// kind: 0 sFFT (DST-I), 1 cFFT (DCT-I), 2 pFFT_1 (periodic forward), 3 pFFT (periodic inverse).
// s and S hold N+1 entries (only the reference's documented range is meaningful).
void ref_fft_d(double* S, const double* s, int N, double dx, int kind) { sfft<double>(S, s, N, dx, kind); }
void ref_fft_f(float* S, const float* s, int N, float dx, int kind) { sfft<float>(S, s, N, dx, kind); }
void ref_sft_d(double* S, const double* s, int N, double dx) {
    std::vector<double> in(s, s + N + 1);
    sFT(S, in.data(), dx, N);
}
void ref_cft_d(double* S, const double* s, int N, double dx) {
    std::vector<double> in(s, s + N + 1);
    cFT(S, in.data(), dx, N);
}
void ref_lapl_cube_d(double* ans, const double* rhs, int periodic, double dx, double dy, double dz,
                     double lx, double ly, double lz, int nx, int ny, int nz) {
    cube<double>(ans, rhs, periodic, dx, dy, dz, lx, ly, lz, nx, ny, nz);
}
void ref_lapl_cube_f(float* ans, const float* rhs, int periodic, double dx, double dy, double dz,
                     double lx, double ly, double lz, int nx, int ny, int nz) {
    cube<float>(ans, rhs, periodic, dx, dy, dz, lx, ly, lz, nx, ny, nz);
}
void* ref_cube_create_d(double dx, double dy, double dz, double lx, double ly, double lz, int nx, int ny,
                        int nz) {
    return new CubeD(dx, dy, dz, lx, ly, lz, nx, ny, nz);
}
void ref_cube_solve_d(void* h, double* ans, double* rhs) { static_cast<CubeHandle*>(h)->solve(ans, rhs); }
void ref_cube_destroy(void* h) { delete static_cast<CubeHandle*>(h); }

void ref_thomas_d(double* B, double* A1, double* A2, double* A3, int n) {
    solve_tdiag_linear_my(B, A1, A2, A3, n);
}
void ref_cr_d(double* d, double* e, double* f, double* b, int q, int n, int general) {
    if (general) fdm::cyclic_reduction_general(d, e, f, b, q, n);
    else fdm::cyclic_reduction(d, e, f, b, q, n);
}
void ref_cr_class_d(double* d, double* e, double* f, double* b, int n) {
    fdm::CyclicReduction<double> cr(n);
    cr.prepare(d, e, f);
    cr.execute(d, e, f, b);
}
}
