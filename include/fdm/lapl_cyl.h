// Drop-in replacement for the reference's LaplCyl3FFT2 (src/lapl_cyl.h:173-249):
// same template parameters, constructor and solve(ans, rhs) on arrays
// [nphi][z1..zn][1..nr] (r fastest); the solve runs on the MI355X through
// include/fdmgpu.h.  The r systems are always solved by cyclic reduction (the
// reference's use_cyclic_reduction = true path); the default LU instantiation
// solves the same systems, so both template values map to it.
//
//   fdm::LaplCyl3FFT2<double, false> lapl(dr, dz, r0, lr, lz, nr, nz, nphi);
//   lapl.solve(ans, rhs);
#pragma once

#include <cmath>

#include "fdmgpu_shim.h"
#include "tensor_flags.h"

namespace fdm {

template <typename T, bool check, tensor_flag zflag = tensor_flag::none, bool use_cyclic_reduction = false>
class LaplCyl3FFT2 {
public:
    // LaplCyl3Data (src/lapl_cyl.h:12-37)
    const double dr, dz, dphi;
    const double dr2, dz2, dphi2;
    const double r0, lr, lz;
    const double slz;
    const int zpoints;
    const int nr, nz, nphi;
    const int nrq;
    const int z1, zn;

    LaplCyl3FFT2(double dr, double dz, double r0, double lr, double lz, int nr, int nz, int nphi)
        : dr(dr), dz(dz), dphi(2 * M_PI / nphi), dr2(dr * dr), dz2(dz * dz), dphi2(dphi * dphi), r0(r0), lr(lr),
          lz(lz), slz(std::sqrt(2. / lz)), zpoints(zflag == tensor_flag::none ? nz + 1 : nz), nr(nr), nz(nz),
          nphi(nphi), nrq((int)std::ceil(std::log2(nr + 1))), z1(zflag == tensor_flag::none ? 1 : 0),
          zn(zflag == tensor_flag::none ? nz : nz - 1) {
        const int prec = sizeof(T) == sizeof(float) ? FDMGPU_F32 : FDMGPU_F64;
        FDMGPU_VERIFY(fdmgpu_plan_cyl3_fft2(&plan_, prec, zflag == tensor_flag::none ? 0 : FDMGPU_PERIODIC_Z, dr, dz,
                                            r0, lr, lz, nr, nz, nphi));
    }

    LaplCyl3FFT2(const LaplCyl3FFT2&) = delete;
    LaplCyl3FFT2& operator=(const LaplCyl3FFT2&) = delete;
    ~LaplCyl3FFT2() {
        if (plan_) fdmgpu_destroy(plan_);
    }

    void solve(T* ans, T* rhs) { FDMGPU_VERIFY(fdmgpu_solve_host(plan_, ans, rhs)); }

    fdmgpu_plan plan() const { return plan_; }

private:
    fdmgpu_plan plan_ = nullptr;
};

}  // namespace fdm
