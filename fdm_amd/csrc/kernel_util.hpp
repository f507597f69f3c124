// Launch helpers shared by the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>
#include <unordered_map>

namespace fdmgpu {

// A kernel that uses more than 64 KB of dynamic LDS (gfx950: 160 KB per CU)
// has to opt in; the opt-in is made once per kernel and size (it is a host
// call on the launch path of every solve otherwise).
template <typename K>
inline hipError_t allow_lds(K kernel, size_t bytes) {
    if (bytes <= 65536) return hipSuccess;
    static std::mutex mu;
    static std::unordered_map<const void*, size_t> done;
    const void* k = reinterpret_cast<const void*>(kernel);
    std::lock_guard<std::mutex> lk(mu);
    auto it = done.find(k);
    if (it != done.end() && it->second >= bytes) return hipSuccess;
    hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e == hipSuccess) done[k] = bytes;
    return e;
}

constexpr size_t kLdsBytes = 160 * 1024;

// Waves per SIMD that the LDS footprint allows for blocks of `threads`
// threads using `lds` bytes each (4 SIMDs per CU), capped at 4: the register
// budget the compiler is held to (512 / waves VGPRs), so that registers never
// cut the occupancy below what LDS allows.
constexpr int lds_waves_per_simd(size_t lds, int threads) {
    const size_t blocks = lds ? kLdsBytes / lds : 8;
    const int waves = (int)(blocks * (size_t)((threads + 63) / 64) / 4);
    return waves < 1 ? 1 : (waves > 4 ? 4 : waves);
}

}  // namespace fdmgpu
