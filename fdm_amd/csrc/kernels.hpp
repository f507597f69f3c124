// Host-side launch interface of the solve kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace fdmgpu {

// Compile-time loop: f(integral_constant<int, i>) for i = B .. E-1.
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// Where problem p of a batch lives: base + p*bs (problems back to back), or
// tab[p] when a device table of per-problem pointers is given (problems in
// separate buffers, e.g. the requests of one serving batch).
template <typename T>
struct BatchPtr {
    T* base;
    const void* const* tab;
    long bs;
    __host__ __device__ T* operator()(int p) const {
        return tab ? (T*)tab[p] : base + (long)p * bs;
    }
    // The same for a wave-uniform p: the table entry comes by a scalar load,
    // counted by lgkmcnt.  A vector load would be counted by vmcnt, and
    // waiting for it would also wait for every vector load and store the
    // wave issued before it (the counter retires in order).
    __device__ T* uniform(int p) const {
        if (!tab) return base + (long)p * bs;
        using CP = const __attribute__((address_space(4))) unsigned long long*;
        return (T*)((CP)tab)[p];
    }
};

// Waits until every vector memory access the wave has issued is complete
// (s_waitcnt vmcnt(0); expcnt and lgkmcnt left alone).
__device__ __forceinline__ void wait_vmem_all() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// A pointer every lane of the wave holds (e.g. read from a per-problem table
// at a wave-uniform index), moved to scalar registers: addresses derived from
// it need no per-lane arithmetic, and a buffer resource built on it needs no
// loop over distinct values.
template <typename T>
__device__ __forceinline__ T* wave_uniform(T* p) {
    const unsigned long long v = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return (T*)(((unsigned long long)hi << 32) | lo);
}

// The same address as a global-memory pointer: loads and stores through it are
// global_* instructions, which cannot alias LDS, so the compiler can keep many
// in flight (through a generic pointer they are flat_* and every one is
// waited for before the next LDS access).
template <typename T>
using global_ptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ global_ptr<T> as_global(T* p) {
    return (global_ptr<T>)p;
}

// A range of global memory addressed through a buffer resource: the byte
// offset of an access is a per-lane VGPR part plus a wave-uniform SGPR part,
// so element offsets that are compile-time constants cost no vector
// instruction (a 64-bit global address costs two per 4 KB window).  Accesses
// at offsets >= `bytes` read 0 and are dropped on store.  The third word
// 0x00020000 is the raw 32-bit data format of the gfx9 family (gfx950).
constexpr int gcd_int(int a, int b) { return b == 0 ? a : gcd_int(b, a % b); }

typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

template <typename T>
struct BufRange {
    __amdgpu_buffer_rsrc_t r;
    __device__ BufRange(const T* p, unsigned bytes)
        : r(__builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)bytes, 0x00020000)) {}
    __device__ __forceinline__ T load(int voff, int soff) const {
        if constexpr (sizeof(T) == 8)
            return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
        else
            return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
    }
    __device__ __forceinline__ void store(T v, int voff, int soff) const {
        if constexpr (sizeof(T) == 8)
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, 0);
        else
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
    }
};

// Pencil transform kinds (see pencil_pairs in dst_device.hpp): DST-I on
// Dirichlet axes, half-complex real DFT forward / inverse on periodic axes.
enum : int { kDst = 0, kRdftFwd = 1, kRdftInv = 2 };

struct PlaneArgs {
    const void* in;
    void* out;
    long bs_in, bs_out;   // batch strides (elements)
    int nplanes, batch;
    double sx, sy;        // DST scale factors for the x and y passes
    const void* tables;   // 3N values of the compute type
    hipStream_t stream;
    bool accum;           // out += result (refinement update) instead of out = result
    const int* done;      // per batch item: skip when non-zero (may be null)
    int kind;             // kDst (0), kRdftFwd (1), kRdftInv (2); periodic kinds need Tin == Tout == T
    const void* const* in_tab;    // per-problem pointers (device) instead of in + p*bs_in, or null
    const void* const* out_tab;
    const float* base = nullptr;  // with accum: out = base + result (base contiguous, stride bs_out);
                                  // problems flagged done get out = base
    bool stream_planes = false;   // plane_stream_kernel: FP64, N = 128, plain (no accum / base / done)
    const void* tables_y = nullptr;   // rectangular planes: the y axis's DST tables (tables: x)
};

struct AxisArgs {
    const void* in;
    void* out;
    long nrows;           // x: rows per batch item (ny*nz)
    int nx, nplanes;      // y: row length and number of planes
    long bs_in, bs_out;
    int batch;
    double scale;
    const void* tables;
    hipStream_t stream;
    bool accum;           // x kernel only
    const int* done;
    int kind;             // as PlaneArgs::kind
    const void* const* in_tab;
    const void* const* out_tab;
    int tile;             // pencils per block (16/32/64; 0 = default_tile)
};

struct ZcrArgs {
    void* data;
    long plane;           // nx*ny modes per z-plane
    int nx;
    const void* lmx;      // eigenvalues per mode index (compute type)
    const void* lmy;
    double b, d0, gamma;  // 1/hz^2, -2/hz^2, rhs scale
    long bs;
    int batch;
    hipStream_t stream;
    const int* done;
    bool periodic;        // circulant system of 2^q rows instead of 2^q - 1 rows
    bool wide;            // float data, double arithmetic
    const void* const* data_tab;
    int gmul;             // lanes-per-system multiplier (1 or 2; 0 = 1)
};

struct RowCrArgs {
    void* data;
    int nx, nrows;
    const void* lmy;      // per row
    const void* sc;       // lm_y_scale per column (compute type)
    const void* Ls;
    const void* Us;
    double inv_dx2;
    long bs;
    int batch;
    hipStream_t stream;
    const int* done;
    bool wide;            // float data, double arithmetic
    const void* const* data_tab;
    const void* lmz = nullptr;   // LaplCyl3FFT2: second per-row eigenvalue, row k uses lmy[k / rdiv] and
    int rdiv = 1;                //   lmz[k % rdiv] (D_j = -2/dx^2 - lmy*s_j - lmz)
    bool pivot = false;          // Gaussian elimination with partial pivoting (rows that are not
                                 //   diagonally dominant) instead of cyclic reduction
};

// LaplRectFFT2 spectral solve: c[k][j] /= -(lmy[k] * sc[j] + lmx[j]) in the
// compute type; with `zero_mode` the (0,0) coefficient becomes 1.
struct Rect2DivArgs {
    void* data;
    int nx, ny;
    const void* lmy;      // per y mode (compute type)
    const void* lmx;      // per x mode
    const void* sc;       // lm_y_scale per x mode
    bool zero_mode;
    long bs;
    int batch;
    hipStream_t stream;
    const int* done;
    const void* const* data_tab;
};

// Forward plane DST of r = f - A u (FP64 residual fused into the load) to FP32,
// with per-plane (sum r^2, sum f^2) partials for the convergence check.
struct PlaneResidArgs {
    const double* f;
    const void* u;        // FP64 iterate
    float* out;
    double* partial;      // [batch][nplanes][2]
    long bs;
    int nplanes, batch;
    double ix, iy, iz;    // 1/h^2 of the inverted operator per axis
    double sx, sy;
    const void* tables;   // FP32 DST tables
    hipStream_t stream;
    const int* done;
    const void* const* f_tab;   // per-problem pointers or null
    const void* const* u_tab;
};

template <typename Tin, typename Tout, typename T>
hipError_t launch_plane(int logn, const PlaneArgs& a);
template <typename Tu>
hipError_t launch_plane_resid(int logn, const PlaneResidArgs& a);
template <typename T>
bool plane_fits(int logn);
// a streaming plane kernel exists for this plane (N = 128: plane_stream_kernel
// for FP64, plane_stream32_kernel for FP32)
bool plane_stream_ok(int logn, bool f64);
// Dirichlet planes of (2^logy - 1) x (2^logx - 1) points, logx != logy, both
// 4 .. 7: the rectangular plane kernel (plane_rect_kernels.hip)
template <typename Tin, typename Tout, typename T>
hipError_t launch_plane_rect(int logx, int logy, const PlaneArgs& a);
bool plane_rect_fits(int logx, int logy, bool f64);
template <typename Tin, typename Tout, typename T>
hipError_t launch_xaxis(int logn, const AxisArgs& a);
template <typename Tin, typename Tout, typename T>
hipError_t launch_yaxis(int logn, const AxisArgs& a);
template <typename T>
hipError_t launch_zcr(int q, const ZcrArgs& a);
template <typename T>
bool tile_ok(int logn, int tile);
template <typename T>
int default_tile(int logn);
bool zcr_gmul_ok(int q, int gm);
template <typename T>
hipError_t launch_rowcr(const RowCrArgs& a);
template <typename T>
hipError_t launch_rect2_divide(const Rect2DivArgs& a);

}  // namespace fdmgpu
