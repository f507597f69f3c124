// Drop-in replacements for the reference's LaplRect (src/lapl_rect.h:11-69)
// and LaplRectFFT2 (:76-109):
// same template parameters, constructor, public coefficient scales and
// solve(ans, rhs); the solve runs on the MI355X through include/fdmgpu.h.
//
//   fdm::LaplRect<double, false> lapl(dx, dy, lx, ly, nx, ny);
//   lapl.lm_y_scale[j] = ...;   // optional, j = 1..nx (cylindrical slices)
//   lapl.solve(ans, rhs);       // host arrays [ny][nx], x fastest
#pragma once

#include <cmath>
#include <vector>

#include "fdmgpu_shim.h"
#include "tensor_flags.h"

namespace fdm {

template <typename T, bool check, typename F = tensor_flags<>>
class LaplRect {
public:
    const double dx, dy;
    const double dx2, dy2;
    const double lx, ly;
    const double slx, sly;
    const int nx, ny;
    const int y1, yn, ypoints;

    // public, settable as in the reference (src/lapl_rect.h:57-59): size nx+1,
    // indexed 1..nx, default 1; pushed to the device before the next solve.
    std::vector<T> lm_y_scale, L_scale, U_scale;

    LaplRect(double dx, double dy, double lx, double ly, int nx, int ny)
        : dx(dx), dy(dy), dx2(dx * dx), dy2(dy * dy), lx(lx), ly(ly), slx(std::sqrt(2. / lx)),
          sly(std::sqrt(2. / ly)), nx(nx), ny(ny), y1(periodic_y() ? 0 : 1), yn(periodic_y() ? ny - 1 : ny),
          ypoints(periodic_y() ? ny : ny + 1), lm_y_scale(nx + 1, T(1)), L_scale(nx + 1, T(1)),
          U_scale(nx + 1, T(1)), pushed_s_(nx + 1, T(1)), pushed_l_(nx + 1, T(1)), pushed_u_(nx + 1, T(1)) {
        FDMGPU_VERIFY(fdmgpu_plan_rect(&plan_, gpu::precision_for<T>(periodic_y()), periodic_y() ? FDMGPU_PERIODIC_Y : 0, dx,
                                       dy, lx, ly, nx, ny));
    }

    LaplRect(const LaplRect&) = delete;
    LaplRect& operator=(const LaplRect&) = delete;
    ~LaplRect() {
        if (plan_) fdmgpu_destroy(plan_);
    }

    // rhs, ans: ypoints-1 (Dirichlet: ny) rows of nx values, x fastest.
    void solve(T* ans, T* rhs) {
        push_scales();
        FDMGPU_VERIFY(fdmgpu_solve_host(plan_, ans, rhs));
    }

    fdmgpu_plan plan() const { return plan_; }

protected:
    static constexpr bool periodic_y() { return has_tensor_flag(F::head, tensor_flag::periodic); }

    void push_scales() {
        if (lm_y_scale == pushed_s_ && L_scale == pushed_l_ && U_scale == pushed_u_) return;
        std::vector<double> s(lm_y_scale.begin(), lm_y_scale.end()), l(L_scale.begin(), L_scale.end()),
            u(U_scale.begin(), U_scale.end());
        FDMGPU_VERIFY(fdmgpu_rect_set_scales(plan_, s.data(), l.data(), u.data()));
        pushed_s_ = lm_y_scale;
        pushed_l_ = L_scale;
        pushed_u_ = U_scale;
    }

    fdmgpu_plan plan_ = nullptr;
    std::vector<T> pushed_s_, pushed_l_, pushed_u_;
};

// Drop-in for LaplRectFFT2 (src/lapl_rect.h:76-109): transforms along both
// axes; F::head is y, F::tail::head is x (tensor_flags<periodic, periodic>
// makes both periodic).  Public members as the reference's (the LaplRect
// ones plus x1/xn/xpoints); lm_y_scale multiplies lm_y per x mode.
template <typename T, bool check, typename F = tensor_flags<>>
class LaplRectFFT2 {
public:
    const double dx, dy;
    const double dx2, dy2;
    const double lx, ly;
    const double slx, sly;
    const int nx, ny;
    const int y1, yn, ypoints;
    const int x1, xn, xpoints;
    std::vector<T> lm_y_scale, L_scale, U_scale;   // L_scale / U_scale do not enter this solver

    LaplRectFFT2(double dx, double dy, double lx, double ly, int nx, int ny)
        : dx(dx), dy(dy), dx2(dx * dx), dy2(dy * dy), lx(lx), ly(ly), slx(std::sqrt(2. / lx)),
          sly(std::sqrt(2. / ly)), nx(nx), ny(ny), y1(periodic_y() ? 0 : 1), yn(periodic_y() ? ny - 1 : ny),
          ypoints(periodic_y() ? ny : ny + 1), x1(periodic_x() ? 0 : 1), xn(periodic_x() ? nx - 1 : nx),
          xpoints(periodic_x() ? nx : nx + 1), lm_y_scale(nx + 1, T(1)), L_scale(nx + 1, T(1)),
          U_scale(nx + 1, T(1)), pushed_s_(nx + 1, T(1)) {
        const int mask = (periodic_y() ? FDMGPU_PERIODIC_Y : 0) | (periodic_x() ? FDMGPU_PERIODIC_X : 0);
        FDMGPU_VERIFY(fdmgpu_plan_rect_fft2(&plan_, gpu::precision_for<T>(mask != 0), mask, dx, dy, lx, ly, nx, ny));
    }

    LaplRectFFT2(const LaplRectFFT2&) = delete;
    LaplRectFFT2& operator=(const LaplRectFFT2&) = delete;
    ~LaplRectFFT2() {
        if (plan_) fdmgpu_destroy(plan_);
    }

    // rhs, ans: rows y1..yn of columns x1..xn, x fastest.
    void solve(T* ans, T* rhs) {
        if (lm_y_scale != pushed_s_) {
            std::vector<double> s(lm_y_scale.begin(), lm_y_scale.end());
            FDMGPU_VERIFY(fdmgpu_rect_set_scales(plan_, s.data(), nullptr, nullptr));
            pushed_s_ = lm_y_scale;
        }
        FDMGPU_VERIFY(fdmgpu_solve_host(plan_, ans, rhs));
    }

    fdmgpu_plan plan() const { return plan_; }

private:
    static constexpr bool periodic_y() { return has_tensor_flag(F::head, tensor_flag::periodic); }
    static constexpr bool periodic_x() { return has_tensor_flag(F::tail::head, tensor_flag::periodic); }

    fdmgpu_plan plan_ = nullptr;
    std::vector<T> pushed_s_;
};

}  // namespace fdm
