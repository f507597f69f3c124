// C++ drop-in check: the reference's unit-test scenarios (ut/ut_lapl_cube.cpp,
// ut/ut_lapl_rect.cpp) written against include/fdm/*.h, i.e. the way an fdm
// caller uses LaplCube / LaplRect, running on the GPU through the C ABI.
//
// Built by __graft_entry__.build() (g++ against libfdmgpu.so); run by
// tests/test_cpp_dropin.py on a GPU.  Exit status 0 = all checks passed.
#define FDMGPU_THROW 1
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <vector>

#include "fdm/lapl_cube.h"
#include "fdm/lapl_rect.h"
#include "fdm/lapl_cyl.h"

static int failures = 0;
#define CHECK(cond, ...)                                  \
    do {                                                  \
        if (!(cond)) {                                    \
            std::printf("FAIL %s:%d ", __FILE__, __LINE__); \
            std::printf(__VA_ARGS__);                     \
            std::printf("\n");                            \
            failures++;                                   \
        }                                                 \
    } while (0)

static double sq(double x) { return x * x; }

// u = sin^2 x + cos^2 y + sin^2 z on the cell-centred grid of [0,1]^3 with the
// boundary values moved to the right-hand side; the discretisation error of
// the 7-point scheme at 31^3 is below 1e-4 (ut/ut_lapl_cube.cpp:39-125).
template <typename T>
double cube_analytic(int n, int mode) {
    fdm::gpu::set_double_mode(mode);
    const double h = 1.0 / n;
    auto ans = [&](int i, int k, int j) {
        double x = h * j - h / 2, y = h * k - h / 2, z = h * i - h / 2;
        return sq(std::sin(x)) + sq(std::cos(y)) + sq(std::sin(z));
    };
    auto rp = [&](int i, int k, int j) {
        double x = h * j - h / 2, y = h * k - h / 2, z = h * i - h / 2;
        return 2 * sq(std::sin(y)) - 2 * sq(std::cos(y)) - 2 * sq(std::sin(x)) + 2 * sq(std::cos(x)) -
               2 * sq(std::sin(z)) + 2 * sq(std::cos(z));
    };
    fdm::LaplCube<T, true> lapl(h, h, h, 1 + h, 1 + h, 1 + h, n, n, n);
    std::vector<T> rhs((size_t)n * n * n), sol(rhs.size());
    for (int i = 1; i <= n; i++)
        for (int k = 1; k <= n; k++)
            for (int j = 1; j <= n; j++) {
                double r = rp(i, k, j);
                if (i == 1) r -= ans(i - 1, k, j) / (h * h);
                if (k == 1) r -= ans(i, k - 1, j) / (h * h);
                if (j == 1) r -= ans(i, k, j - 1) / (h * h);
                if (j == n) r -= ans(i, k, j + 1) / (h * h);
                if (k == n) r -= ans(i, k + 1, j) / (h * h);
                if (i == n) r -= ans(i + 1, k, j) / (h * h);
                rhs[((size_t)(i - 1) * n + (k - 1)) * n + (j - 1)] = (T)r;
            }
    lapl.solve(sol.data(), rhs.data());
    double err = 0, nrm = 0;
    for (int i = 1; i <= n; i++)
        for (int k = 1; k <= n; k++)
            for (int j = 1; j <= n; j++) {
                double f = ans(i, k, j);
                err = std::max(err, std::fabs(sol[((size_t)(i - 1) * n + (k - 1)) * n + (j - 1)] - f));
                nrm = std::max(nrm, std::fabs(f));
            }
    return err / nrm;
}

// biquadratic u = (x-x1)(x-x2)(y-y1)(y-y2): the 5-point scheme is exact
// (ut/ut_lapl_rect.cpp:300-374): < 1e-14 double, < 1e-5 float.
template <typename T>
double rect_biquadratic(int n) {
    const double h = 1.0 / n;
    auto a0 = [&](int k, int j) {
        double x = h * j - h / 2, y = h * k - h / 2;
        return x * (x - 1) * y * (y - 1);
    };
    fdm::LaplRect<T, true> lapl(h, h, 1 + h, 1 + h, n, n);
    std::vector<T> rhs((size_t)n * n), sol(rhs.size());
    for (int k = 1; k <= n; k++)
        for (int j = 1; j <= n; j++) {
            double x = h * j - h / 2, y = h * k - h / 2;
            double r = 2 * y * (y - 1) + 2 * x * (x - 1);
            if (k == 1) r -= a0(k - 1, j) / (h * h);
            if (j == 1) r -= a0(k, j - 1) / (h * h);
            if (j == n) r -= a0(k, j + 1) / (h * h);
            if (k == n) r -= a0(k + 1, j) / (h * h);
            rhs[(size_t)(k - 1) * n + (j - 1)] = (T)r;
        }
    lapl.solve(sol.data(), rhs.data());
    double err = 0, nrm = 0;
    for (int k = 1; k <= n; k++)
        for (int j = 1; j <= n; j++) {
            err = std::max(err, std::fabs(sol[(size_t)(k - 1) * n + (j - 1)] - a0(k, j)));
            nrm = std::max(nrm, std::fabs(a0(k, j)));
        }
    return err / nrm;
}

// LaplRect (transform + tridiagonal) against LaplRectFFT2 (two transforms) on
// the analytic rhs of ut/ut_lapl_rect.cpp:384-456 at 511^2; the reference
// asserts agreement to 1e-15 absolute (its answers are O(1e-1)).
double rect_fft1_fft2(int n) {
    const double h = 1.0 / n;
    fdm::LaplRect<double, true> lapl1(h, h, 1 + h, 1 + h, n, n);
    fdm::LaplRectFFT2<double, true> lapl2(h, h, 1 + h, 1 + h, n, n);
    std::vector<double> rhs((size_t)n * n), a1(rhs.size()), a2(rhs.size());
    for (int k = 1; k <= n; k++)
        for (int j = 1; j <= n; j++) {
            const double x = h * j - h / 2, y = h * k - h / 2;
            rhs[(size_t)(k - 1) * n + (j - 1)] = 2 * std::sin(y) * std::sin(y) - 2 * std::cos(y) * std::cos(y) -
                                                 2 * std::sin(x) * std::sin(x) + 2 * std::cos(x) * std::cos(x);
        }
    lapl1.solve(a1.data(), rhs.data());
    lapl2.solve(a2.data(), rhs.data());
    double err = 0;
    for (size_t i = 0; i < a1.size(); i++) err = std::max(err, std::fabs(a1[i] - a2[i]));
    return err;
}

// ut/ut_lapl_cyl.cpp:320-406: second-order convergence of LaplCyl3FFT2 on
// u = (z-h0)(z-h1)(r-r0)(r-R)(sin phi + cos phi) (cell-centred r, z nodes;
// rhs = the continuous Laplacian P[Z(4r-r0-R)/r + 2Rr - Rr Z/r^2], boundary
// values moved to the rhs).  Returns max|u_h - u| / max|u|.
double cyl_known_answer(int nr, int nz, int nphi) {
    const double r0 = M_PI / 2, R = M_PI, h0 = 0, h1 = 10;
    const double dr = (R - r0) / nr, dz = (h1 - h0) / nz, dphi = 2 * M_PI / nphi;
    auto u = [&](int i, int k, int j) {
        const double r = r0 + dr * j - dr / 2, z = h0 + dz * k - dz / 2, phi = dphi * (i + 1) - dphi / 2;
        return (z - h0) * (z - h1) * (r - r0) * (r - R) * (std::sin(phi) + std::cos(phi));
    };
    fdm::LaplCyl3FFT2<double, true> lapl(dr, dz, r0 - dr / 2, R - r0 + dr, h1 - h0 + dz, nr, nz, nphi);
    std::vector<double> rhs((size_t)nphi * nz * nr), sol(rhs.size());
    for (int i = 0; i < nphi; i++)
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= nr; j++) {
                const double r = r0 + dr * j - dr / 2, z = h0 + dz * k - dz / 2, phi = dphi * (i + 1) - dphi / 2;
                const double Z = (z - h0) * (z - h1), Rr = (r - r0) * (r - R), P = std::sin(phi) + std::cos(phi);
                double f = P * (Z * (4 * r - r0 - R) / r + 2 * Rr - Rr * Z / (r * r));
                if (j == 1) f -= (r - dr / 2) / r * u(i, k, 0) / (dr * dr);
                if (j == nr) f -= (r + dr / 2) / r * u(i, k, nr + 1) / (dr * dr);
                if (k == 1) f -= u(i, 0, j) / (dz * dz);
                if (k == nz) f -= u(i, nz + 1, j) / (dz * dz);
                rhs[((size_t)i * nz + (k - 1)) * nr + (j - 1)] = f;
            }
    lapl.solve(sol.data(), rhs.data());
    double err = 0, nrm = 0;
    for (int i = 0; i < nphi; i++)
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= nr; j++) {
                err = std::max(err, std::fabs(sol[((size_t)i * nz + (k - 1)) * nr + (j - 1)] - u(i, k, j)));
                nrm = std::max(nrm, std::fabs(u(i, k, j)));
            }
    return err / nrm;
}

// periodic^3 on [0, 2pi)^3, n = 32: u = sin^3 x + cos^5 y - sin z with its
// analytic Laplacian as rhs; the solve's mean is removed and the max error
// must stay below 1e-2 (ut/ut_lapl_cube.cpp:134-244).
template <typename T>
double cube_periodic(int n) {
    using flags = fdm::tensor_flags<fdm::tensor_flag::periodic, fdm::tensor_flag::periodic, fdm::tensor_flag::periodic>;
    const double h = 2 * M_PI / n;
    fdm::LaplCube<T, true, flags> lapl(h, h, h, 2 * M_PI, 2 * M_PI, 2 * M_PI, n, n, n);
    auto coord = [&](int i) { return h * i - h / 2; };
    std::vector<T> rhs((size_t)n * n * n), sol(rhs.size());
    for (int i = 0; i < n; i++)
        for (int k = 0; k < n; k++)
            for (int j = 0; j < n; j++) {
                const double x = coord(j), y = coord(k), z = coord(i);
                rhs[((size_t)i * n + k) * n + j] =
                    (T)(1. / 16. * (-12. * std::sin(x) + 36. * std::sin(3 * x) - 10. * std::cos(y) -
                                    45. * std::cos(3 * y) - 25. * std::cos(5 * y) + 16. * std::sin(z)));
            }
    lapl.solve(sol.data(), rhs.data());
    double avg = 0;
    for (T v : sol) avg += v;
    avg /= (double)sol.size();
    double err = 0, nrm = 0;
    for (int i = 0; i < n; i++)
        for (int k = 0; k < n; k++)
            for (int j = 0; j < n; j++) {
                const double x = coord(j), y = coord(k), z = coord(i);
                const double f = std::pow(std::sin(x), 3) + std::pow(std::cos(y), 5) - std::sin(z);
                err = std::max(err, std::fabs(sol[((size_t)i * n + k) * n + j] - avg - f));
                nrm = std::max(nrm, std::fabs(f));
            }
    return err / nrm;
}

int main() {
    double e;
    e = cube_analytic<double>(31, FDMGPU_F64);
    std::printf("LaplCube<double> 31^3 analytic (f64): err %.3e\n", e);
    CHECK(e < 1e-4, "cube double err %g", e);
    e = cube_analytic<double>(31, FDMGPU_F32_REFINE64);
    std::printf("LaplCube<double> 31^3 analytic (refine): err %.3e\n", e);
    CHECK(e < 1e-4, "cube double refine err %g", e);
    e = cube_analytic<float>(31, FDMGPU_F64);
    std::printf("LaplCube<float> 31^3 analytic: err %.3e\n", e);
    CHECK(e < 1e-4, "cube float err %g", e);
    fdm::gpu::set_double_mode(FDMGPU_F64);
    e = cube_periodic<double>(32);
    std::printf("LaplCube<double, periodic^3> 32^3 analytic: err %.3e\n", e);
    CHECK(e < 1e-2, "cube periodic double err %g", e);
    fdm::gpu::set_double_mode(FDMGPU_F32_REFINE64);
    const double er = cube_periodic<double>(32);
    std::printf("LaplCube<double, periodic^3> 32^3 analytic (refine): err %.3e\n", er);
    CHECK(std::fabs(er - e) <= 1e-9 * (1 + e), "cube periodic refine err %g vs FP64 %g", er, e);
    fdm::gpu::set_double_mode(FDMGPU_F64);
    e = cube_periodic<float>(32);
    std::printf("LaplCube<float, periodic^3> 32^3 analytic: err %.3e\n", e);
    CHECK(e < 1e-2, "cube periodic float err %g", e);
    e = rect_biquadratic<double>(31);
    std::printf("LaplRect<double> 31^2 biquadratic: err %.3e\n", e);
    CHECK(e < 1e-14, "rect double err %g", e);
    e = rect_biquadratic<float>(31);
    std::printf("LaplRect<float> 31^2 biquadratic: err %.3e\n", e);
    CHECK(e < 1e-5, "rect float err %g", e);

    {
        const double e1 = cyl_known_answer(16, 15, 16), e2 = cyl_known_answer(32, 31, 32);
        std::printf("LaplCyl3FFT2<double> known answer: err %.3e (16,15,16), %.3e (32,31,32), ratio %.2f\n", e1, e2,
                    e1 / e2);
        CHECK(e1 / e2 > 3.7, "cyl convergence ratio %g", e1 / e2);
    }
    e = rect_fft1_fft2(511);
    std::printf("LaplRect vs LaplRectFFT2<double> 511^2: max abs diff %.3e\n", e);
    CHECK(e < 1e-13, "rect fft1/fft2 diff %g", e);

    // the reference aborts for a non-power-of-two transform length (src/fft.cpp:67);
    // with FDMGPU_THROW the shim throws instead
    bool threw = false;
    try {
        fdm::LaplCube<double, false> bad(0.1, 0.1, 0.1, 1, 1, 1, 6, 7, 7);
    } catch (const std::runtime_error&) {
        threw = true;
    }
    CHECK(threw, "non-power-of-two size must fail");

    std::printf("%s (%d failures)\n", failures ? "FAILED" : "OK", failures);
    return failures ? 1 : 0;
}
