// Shared plumbing of the header-only drop-in classes: error policy and the
// choice of GPU precision for double solvers.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

#include "../fdmgpu.h"

namespace fdm {
namespace gpu {

// How LaplCube<double>/LaplRect<double> solve on the GPU: plain FP64
// (FDMGPU_F64, default) or FP32 transforms with FP64 iterative refinement
// (FDMGPU_F32_REFINE64).  Environment override: FDMGPU_DOUBLE_MODE=refine|f64.
inline int& double_mode_ref() {
    static int mode = [] {
        const char* e = std::getenv("FDMGPU_DOUBLE_MODE");
        return (e && std::strcmp(e, "refine") == 0) ? FDMGPU_F32_REFINE64 : FDMGPU_F64;
    }();
    return mode;
}
inline void set_double_mode(int precision) { double_mode_ref() = precision; }

// Periodic 2-D solvers (LaplRect, LaplRectFFT2) always solve in plain FP64:
// the projected-residual refinement of a singular operator exists for
// periodic cubes only (refine_periodic = true).
template <typename T>
inline int precision_for(bool periodic = false, bool refine_periodic = false) {
    if (sizeof(T) == sizeof(float)) return FDMGPU_F32;
    return periodic && !refine_periodic ? FDMGPU_F64 : double_mode_ref();
}

// The reference reports failures through verify(): message on stderr, then
// abort() (src/verify.h:1-36).  Define FDMGPU_THROW to get exceptions instead.
inline void check(int rc, const char* what, const char* file, int line) {
    if (rc == FDMGPU_OK) return;
#ifdef FDMGPU_THROW
    throw std::runtime_error(std::string(what) + ": " + fdmgpu_strerror(rc));
#else
    std::fprintf(stderr, "verify(%s) failed at %s:%d: %s (fdmgpu code %d)\n", what, file, line,
                 fdmgpu_strerror(rc), rc);
    std::abort();
#endif
}

}  // namespace gpu
}  // namespace fdm

#define FDMGPU_VERIFY(expr) ::fdm::gpu::check((expr), #expr, __FILE__, __LINE__)
