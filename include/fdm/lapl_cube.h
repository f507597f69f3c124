// Drop-in replacement for the reference's LaplCube (src/lapl_cube.h:9-106):
// same template parameters, constructor and solve(ans, rhs); the solve runs
// on the MI355X through the C ABI in include/fdmgpu.h.
//
//   fdm::LaplCube<double, false> lapl(dx, dy, dz, lx, ly, lz, nx, ny, nz);
//   lapl.solve(ans, rhs);     // host arrays [nz][ny][nx], x fastest
//
// `check` only adds index checks to the reference's tensor accessors; the
// GPU path does not index through tensors, so it has no effect here.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>

#include "fdmgpu_shim.h"
#include "tensor_flags.h"

namespace fdm {

template <typename T, bool check, typename F = tensor_flags<>>
class LaplCube {
public:
    static constexpr tensor_flag zflag = F::head;
    static constexpr tensor_flag yflag = F::tail::head;
    static constexpr tensor_flag xflag = F::tail::tail::head;

    const double dx, dy, dz;
    const double dx2, dy2, dz2;
    const double lx, ly, lz;
    const double slx, sly, slz;

    // index bounds as the reference computes them (x uses the z flag,
    // src/lapl_cube.h:68,72,76)
    const int z1, y1, x1;
    const int zn, yn, xn;
    const int zpoints, ypoints, xpoints;

    const int nx, ny, nz;
    const int mxdim;
    const std::array<int, 6> indices;

    LaplCube(double dx, double dy, double dz, double lx, double ly, double lz, int nx, int ny, int nz)
        : dx(dx), dy(dy), dz(dz), dx2(dx * dx), dy2(dy * dy), dz2(dz * dz), lx(lx), ly(ly), lz(lz),
          slx(std::sqrt(2. / lx)), sly(std::sqrt(2. / ly)), slz(std::sqrt(2. / lz)),
          z1(periodic_z() ? 0 : 1), y1(periodic_y() ? 0 : 1), x1(periodic_z() ? 0 : 1),
          zn(periodic_z() ? nz - 1 : nz), yn(periodic_y() ? ny - 1 : ny), xn(periodic_z() ? nx - 1 : nx),
          zpoints(periodic_z() ? nz : nz + 1), ypoints(periodic_y() ? ny : ny + 1),
          xpoints(periodic_z() ? nx : nx + 1), nx(nx), ny(ny), nz(nz),
          mxdim(std::max({nx + 1, ny + 1, nz + 1})), indices({1, nz, 1, ny, 1, nx}) {
        const int mask = (has_tensor_flag(xflag, tensor_flag::periodic) ? FDMGPU_PERIODIC_X : 0) |
                         (has_tensor_flag(yflag, tensor_flag::periodic) ? FDMGPU_PERIODIC_Y : 0) |
                         (has_tensor_flag(zflag, tensor_flag::periodic) ? FDMGPU_PERIODIC_Z : 0);
        FDMGPU_VERIFY(fdmgpu_plan_cube(&plan_, gpu::precision_for<T>(mask != 0, true), mask, dx, dy, dz, lx, ly, lz, nx, ny, nz));
    }

    LaplCube(const LaplCube&) = delete;
    LaplCube& operator=(const LaplCube&) = delete;
    ~LaplCube() {
        if (plan_) fdmgpu_destroy(plan_);
    }

    // rhs, ans: nz*ny*nx interior values, x fastest (&RHS[1][1][1] in the
    // reference's tensors); ans is fully overwritten; ans == rhs is allowed.
    void solve(T* ans, T* rhs) { FDMGPU_VERIFY(fdmgpu_solve_host(plan_, ans, rhs)); }

    // Same solve on device pointers, ordered on a hipStream_t (nullptr = default).
    void solve_device(T* d_ans, const T* d_rhs, void* stream = nullptr) {
        FDMGPU_VERIFY(fdmgpu_solve_device(plan_, d_ans, d_rhs, stream));
    }

    fdmgpu_plan plan() const { return plan_; }

private:
    static constexpr bool periodic_z() { return has_tensor_flag(F::head, tensor_flag::periodic); }
    static constexpr bool periodic_y() { return has_tensor_flag(F::tail::head, tensor_flag::periodic); }
    fdmgpu_plan plan_ = nullptr;
};

}  // namespace fdm
