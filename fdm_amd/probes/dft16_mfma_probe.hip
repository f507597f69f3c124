// Measurement probe: the 16-point sub-DFTs of the FP32 plane transforms
// (step 1 of the 4-step N = 128 FFT in dst_device.hpp: one 16-point complex
// DFT per lane, inputs from LDS, outputs to LDS) on the VALU, as the plane
// kernels do, against the same DFTs as a matrix product on the matrix cores
// (v_mfma_f32_32x32x2_f32).  With the 16 complex inputs of a sequence as 32
// reals x = (Re, Im), the DFT is y = M x with the real 32 x 32 matrix
// M = [[Wr, -Wi], [Wi, Wr]], W_jk = exp(-2 pi i jk / 16).  A wavefront
// transforms 64 sequences per round, as two 32-sequence products
// Y^T (32 x 32) = X^T (32 x 32) M^T, each 16 MFMA steps of K = 2: the A
// operand (lane l: X^T[l % 32][2t + l / 32]) and the B operand (lane l:
// M^T[2t + l / 32][l % 32], constant) come from LDS and registers, the C
// result (lane l: column l % 32, rows 8j + 4(l / 32) + i) goes back to LDS.
//
// Each block repeats ROUNDS rounds on its sequences in LDS, so the timing is
// that of the arithmetic and the LDS traffic, not of HBM.  The program checks
// the two results against each other and prints one JSON line.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <vector>

#include "../csrc/dst_device.hpp"

#define CK(x)                                                                                   \
    do {                                                                                        \
        hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess) {                                                                 \
            std::printf("{\"error\": \"%s at line %d\"}\n", hipGetErrorString(e_), __LINE__);   \
            return 1;                                                                           \
        }                                                                                       \
    } while (0)

namespace {

constexpr int kSeq = 64;        // sequences per wavefront (one block = one wavefront)
constexpr int kRounds = 256;    // DFT rounds per launch on the same LDS data
constexpr float kNorm = 0.25f;  // |W| = 4: keep magnitudes bounded round to round

// VALU: lane l transforms sequence l (16 complex values, slots re[0..15], im[0..15])
__global__ __launch_bounds__(64) void dft16_valu(const float* __restrict__ in, float* __restrict__ out) {
    __shared__ float s[kSeq * 33];   // sequence l at s[l * 33 ...] (odd pitch: no bank conflicts)
    const int l = threadIdx.x;
    const float* src = in + (long)blockIdx.x * kSeq * 32;
    for (int i = l; i < kSeq * 32; i += 64) s[(i / 32) * 33 + i % 32] = src[i];
    __syncthreads();
    for (int r = 0; r < kRounds; ++r) {
        float re[16], im[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            re[k] = s[l * 33 + k];
            im[k] = s[l * 33 + 16 + k];
        }
        fdmgpu::dft<16>(re, im);
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            s[l * 33 + k] = re[k] * kNorm;
            s[l * 33 + 16 + k] = im[k] * kNorm;
        }
    }
    __syncthreads();
    float* dst = out + (long)blockIdx.x * kSeq * 32;
    for (int i = l; i < kSeq * 32; i += 64) dst[i] = s[(i / 32) * 33 + i % 32];
}

typedef float f32x16 __attribute__((ext_vector_type(16)));

// MFMA: two products of 32 sequences each per round
__global__ __launch_bounds__(64) void dft16_mfma(const float* __restrict__ in, float* __restrict__ out,
                                                 const float* __restrict__ mt) {
    __shared__ float s[kSeq * 33];
    const int l = threadIdx.x;
    const float* src = in + (long)blockIdx.x * kSeq * 32;
    for (int i = l; i < kSeq * 32; i += 64) s[(i / 32) * 33 + i % 32] = src[i];
    // B operand per step t: M^T[2t + l/32][l % 32] = M[l % 32][2t + l/32], scaled
    float b[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) b[t] = mt[(2 * t + l / 32) * 32 + l % 32] * kNorm;
    __syncthreads();
    const int row = l % 32, kk = l / 32;
    for (int r = 0; r < kRounds; ++r) {
#pragma unroll
        for (int g = 0; g < 2; ++g) {
            f32x16 c = {};
            const int seq = 32 * g + row;
#pragma unroll
            for (int t = 0; t < 16; ++t) c = __builtin_amdgcn_mfma_f32_32x32x2f32(s[seq * 33 + 2 * t + kk], b[t], c, 0, 0, 0);
            __builtin_amdgcn_s_waitcnt(0xc07f);   // lgkmcnt(0): this group's LDS reads are done
            __builtin_amdgcn_wave_barrier();
            // C: lane l holds output n = l % 32 of sequences 8j + 4(l / 32) + i
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int i = 0; i < 4; ++i) s[(32 * g + 8 * j + 4 * kk + i) * 33 + row] = c[4 * j + i];
            __builtin_amdgcn_wave_barrier();
        }
    }
    __syncthreads();
    float* dst = out + (long)blockIdx.x * kSeq * 32;
    for (int i = l; i < kSeq * 32; i += 64) dst[i] = s[(i / 32) * 33 + i % 32];
}

}  // namespace

int main() {
    const int blocks = 256 * 32;   // wavefronts: 32 per compute unit
    const long n = (long)blocks * kSeq * 32;
    std::vector<float> h(n), ho(n), hm(n);
    unsigned long long z = 12345;
    for (long i = 0; i < n; ++i) {
        z = z * 6364136223846793005ull + 1442695040888963407ull;
        h[i] = (float)((z >> 40) * (1.0 / 16777216.0)) - 0.5f;
    }
    // M = [[Wr, -Wi], [Wi, Wr]]: outputs 0..15 real parts, 16..31 imaginary parts
    std::vector<float> M(32 * 32);
    for (int j = 0; j < 16; ++j)
        for (int k = 0; k < 16; ++k) {
            const double a = -2 * M_PI * j * k / 16;
            M[j * 32 + k] = (float)std::cos(a);
            M[j * 32 + 16 + k] = (float)-std::sin(a);
            M[(16 + j) * 32 + k] = (float)std::sin(a);
            M[(16 + j) * 32 + 16 + k] = (float)std::cos(a);
        }
    float *din, *dv, *dm, *dM;
    CK(hipMalloc(&din, n * 4));
    CK(hipMalloc(&dv, n * 4));
    CK(hipMalloc(&dm, n * 4));
    CK(hipMalloc(&dM, M.size() * 4));
    CK(hipMemcpy(din, h.data(), n * 4, hipMemcpyHostToDevice));
    CK(hipMemcpy(dM, M.data(), M.size() * 4, hipMemcpyHostToDevice));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    float ms_v = 0, ms_m = 0;
    for (int rep = 0; rep < 3; ++rep) {   // the last repetition is reported
        CK(hipEventRecord(e0));
        hipLaunchKernelGGL(dft16_valu, dim3(blocks), dim3(64), 0, 0, din, dv);
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        CK(hipEventElapsedTime(&ms_v, e0, e1));
        CK(hipEventRecord(e0));
        hipLaunchKernelGGL(dft16_mfma, dim3(blocks), dim3(64), 0, 0, din, dm, dM);
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        CK(hipEventElapsedTime(&ms_m, e0, e1));
    }
    CK(hipMemcpy(ho.data(), dv, n * 4, hipMemcpyDeviceToHost));
    CK(hipMemcpy(hm.data(), dm, n * 4, hipMemcpyDeviceToHost));
    double num = 0, den = 0;
    for (long i = 0; i < n; ++i) {
        num += (double)(ho[i] - hm[i]) * (ho[i] - hm[i]);
        den += (double)ho[i] * ho[i];
    }
    const double dfts = (double)blocks * kSeq * kRounds;
    std::printf("{\"probe\": \"16-point complex DFTs, FP32, LDS in/out, %d rounds\", \"dfts\": %.0f, "
                "\"valu_ms\": %.4f, \"mfma_ms\": %.4f, \"valu_ns_per_dft\": %.4f, \"mfma_ns_per_dft\": %.4f, "
                "\"mfma_over_valu_time\": %.3f, \"rel_l2_diff\": %.3e}\n",
                kRounds, dfts, ms_v, ms_m, 1e6 * ms_v / dfts, 1e6 * ms_m / dfts, ms_m / ms_v, std::sqrt(num / den));
    return 0;
}
