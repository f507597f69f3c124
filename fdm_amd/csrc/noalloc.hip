// fdmgpu_solve_batch_noalloc: the batched device solve without any device
// allocation, for callers that cannot allocate (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The FP32+refine scratch that fdmgpu_solve_batch_device keeps in the plan
// (allocated on first use) comes from `scratch`, a static device array of
// kNoAllocScratchBytes that every GPU the library runs on holds until the
// process ends.  Host code never passes the array's name to a kernel: the
// device address of each GPU's copy is looked up once per device with
// hipGetSymbolAddress and kept in g_scratch[device].  Every call on one GPU
// shares that GPU's array (calls that overlap on one GPU race: the caller
// orders them, e.g. on one stream), and it keeps the previous call's
// contents; what must start at zero (the per-problem convergence flags and
// iteration counts) is cleared by a fill kernel at the start of every solve.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "plan.hpp"

namespace {

constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;   // 64 MiB per GPU
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

// Device address of this device's copy of `scratch` (the current device is
// the plan's); looked up on the first call on each device.
int scratch_of(int device, void** out) {
    if (device < 0 || device >= kMaxDevices) return FDMGPU_ERR_DEVICE;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            char msg[160];
            std::snprintf(msg, sizeof msg, "hipGetSymbolAddress(scratch) failed on device %d: %s", device,
                          hipGetErrorString(e));
            fdmgpu::set_error_text(msg);
            return FDMGPU_ERR_DEVICE;
        }
        g_scratch[device] = ptr;
    }
    *out = g_scratch[device];
    return FDMGPU_OK;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

extern "C" int fdmgpu_solve_batch_noalloc(fdmgpu_plan P, int count, void* d_ans, const void* d_rhs, void* stream) {
    if (!P || count < 1 || !d_ans || !d_rhs) return FDMGPU_ERR_INVALID_ARG;
    const bool refine = P->precision == FDMGPU_F32_REFINE64;
    const size_t elems = (size_t)fdmgpu::plan_elems(P);
    // scratch: the refinement Work, plus a copy of f when the solve runs in place
    const size_t work = fdmgpu::work_bytes(P, count);
    const size_t fcopy = refine && d_ans == d_rhs ? align256(sizeof(double) * elems * count) : 0;
    if (work + fcopy > kNoAllocScratchBytes) {
        char msg[200];
        std::snprintf(msg, sizeof msg,
                      "fdmgpu_solve_batch_noalloc: %d problem(s) need %zu bytes of scratch, more than the "
                      "%zu-byte static scratch",
                      count, work + fcopy, kNoAllocScratchBytes);
        fdmgpu::set_error_text(msg);
        return FDMGPU_ERR_TOO_LARGE;
    }
    if (hipSetDevice(P->device) != hipSuccess) return FDMGPU_ERR_DEVICE;
    void* mem = nullptr;
    int rc = scratch_of(P->device, &mem);
    if (rc) return rc;
    fdmgpu::Work W = fdmgpu::work_carve(P, count, mem);
    if (fcopy) {
        W.fcopy = (double*)((char*)mem + align256(work));
        W.fcopy_bytes = fcopy;
    }
    // no timing events here (creating them is a runtime allocation)
    const unsigned timing = P->timing;
    P->timing = 0;
    const hipStream_t s = (hipStream_t)stream;
    rc = fdmgpu::plan_solve(P, W, d_ans, nullptr, d_rhs, nullptr, count, s);
    P->timing = timing;
    P->last_stream = s;
    P->last_count = count;
    P->last_iters = W.iters;
    P->last_resid = W.resid;
    return rc;
}
