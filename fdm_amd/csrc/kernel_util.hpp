// Launch helpers shared by the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace fdmgpu {

// A kernel that uses more than 64 KB of dynamic LDS (gfx950: 160 KB per CU)
// has to opt in.
template <typename K>
inline hipError_t allow_lds(K kernel, size_t bytes) {
    if (bytes <= 65536) return hipSuccess;
    return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes);
}

constexpr size_t kLdsBytes = 160 * 1024;

}  // namespace fdmgpu
