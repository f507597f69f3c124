// Device building blocks for the fast Poisson solve on gfx950 (CDNA4, wave64).
//
//  * dft<R>:        forward DFT of R <= 32 complex values held in registers
//                   (fully unrolled radix-2, compile-time twiddles).
//  * dst_pairs<>:   DST-I of pencils held in LDS, two pencils per job.
//
// DST-I definition (the transform FFT<T>::sFFT computes, reference
// src/fft.cpp:293-365):  X[k] = scale * sum_{j=1}^{N-1} x[j] sin(pi j k / N),
// k = 1..N-1, N = 2^LOGN.  Algorithm (not the reference's Samarskii-Nikolaev
// recursion, which is sequential per pencil):
//   1. w_j = sin(pi j/N)(x_j + x_{N-j}) + (x_j - x_{N-j})/2        (w_0 = 0)
//          = (sin(pi j/N) + 1/2)(x_j + x_{N-j}) - x_{N-j}   (one add, one fma)
//      Two pencils a, b are packed as z_j = w^a_j + i w^b_j.
//   2. Z = FFT_N(z), done as a 4-step FFT N = N1 x N2: N1 lanes per job each
//      run an N2-point DFT in registers, twiddle by w_N^{j1 k2}, exchange
//      through the pencils' own LDS slots, then N1-point DFTs.
//   3. With W = the real DFT of w:  Re W_k = X_{2k+1} - X_{2k-1},
//      Im W_k = -X_{2k}.  So X_{2k} comes directly and the odd outputs are a
//      prefix sum X_{2k+1} = Re W_0 / 2 + sum_{i=1..k} Re W_i, computed as a
//      per-lane running sum plus a shuffle scan across the job's N1 lanes.
// Everything happens in place in the pencils' N LDS slots (slot 0 is a
// dummy that reads as zero after the transform), so an N x N plane (one
// dummy row, one dummy column) holds a full 127^2 FP64 plane in 139 KB.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "kernels.hpp"   // kDst, kRdftFwd, kRdftInv

namespace fdmgpu {

constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v >> 1); }
constexpr int bitrev(int v, int bits) {
    int r = 0;
    for (int b = 0; b < bits; b++)
        if (v & (1 << b)) r |= 1 << (bits - 1 - b);
    return r;
}

// cos(2 pi m / 32) from the first quadrant by symmetry (exact zeros).
constexpr double kQuarter[9] = {1.0,
                                0.98078528040323044912618223613424,
                                0.92387953251128675612818318939679,
                                0.83146961230254523707878837761791,
                                0.70710678118654752440084436210485,
                                0.55557023301960222474283081394853,
                                0.38268343236508977172845998403040,
                                0.19509032201612826784828486847702,
                                0.0};
constexpr double cos32(int m) {
    m = ((m % 32) + 32) % 32;
    return m <= 8 ? kQuarter[m] : m <= 16 ? -kQuarter[16 - m] : m <= 24 ? -kQuarter[m - 16] : kQuarter[32 - m];
}
constexpr double sin32(int m) { return cos32(m - 8); }

// (re, im) *= exp(-2 pi i K / L), L | 32, all compile time.
template <int L, int K, typename T>
__device__ __forceinline__ void twiddle_ct(T& re, T& im) {
    constexpr int k = ((K % L) + L) % L;
    if constexpr (k == 0) {
    } else if constexpr (4 * k == L) {   // * (-i)
        T t = re;
        re = im;
        im = -t;
    } else if constexpr (2 * k == L) {
        re = -re;
        im = -im;
    } else if constexpr (4 * k == 3 * L) {   // * (+i)
        T t = re;
        re = -im;
        im = t;
    } else {
        constexpr int m = k * (32 / L);
        constexpr T c = (T)cos32(m);
        constexpr T s = (T)(-sin32(m));
        T r = re * c - im * s;
        T i = re * s + im * c;
        re = r;
        im = i;
    }
}

// In-register forward DFT of size R (natural order in and out).
template <int R, typename T>
__device__ __forceinline__ void dft(T (&re)[R], T (&im)[R]) {
    static_assert(R >= 1 && R <= 32 && (R & (R - 1)) == 0, "radix");
    if constexpr (R > 1) {
        constexpr int LOG = ilog2(R);
        static_for<0, R>([&](auto ii) {
            constexpr int i = decltype(ii)::value;
            constexpr int r = bitrev(i, LOG);
            if constexpr (i < r) {
                T t = re[i];
                re[i] = re[r];
                re[r] = t;
                t = im[i];
                im[i] = im[r];
                im[r] = t;
            }
        });
        static_for<1, LOG + 1>([&](auto ss) {
            constexpr int len = 1 << decltype(ss)::value;
            static_for<0, R / len>([&](auto bb) {
                static_for<0, len / 2>([&](auto kk) {
                    constexpr int a = decltype(bb)::value * len + decltype(kk)::value;
                    constexpr int b = a + len / 2;
                    T tr = re[b], ti = im[b];
                    twiddle_ct<len, decltype(kk)::value>(tr, ti);
                    re[b] = re[a] - tr;
                    im[b] = im[a] - ti;
                    re[a] = re[a] + tr;
                    im[a] = im[a] + ti;
                });
            });
        });
    }
}

template <int LOGN>
struct DstShape {
    static constexpr int N = 1 << LOGN;
    static constexpr int N1 = 1 << (LOGN / 2);   // lanes per job
    static constexpr int N2 = N / N1;            // N2 >= N1
    static constexpr int M = N / 2;
    static constexpr int KC = M / N1;            // post-processing outputs per lane (>= 1)
    static constexpr int LOGKC = ilog2(KC);
    // LDS padding: a pencil-to-pencil stride congruent to N1 mod 32 banks spreads
    // the N1 lanes of neighbouring jobs over distinct banks.
    static constexpr int PAD = N1;

    // Slot of the step-1 result Y[j1][k2] (XOR swizzle: conflict-free for the
    // step-1 stores, contiguous in j1, and the step-2 loads, contiguous in k2).
    __device__ static constexpr int posY(int j1, int k2) { return N1 * k2 + (j1 ^ (k2 & (N1 - 1))); }
    // Slot of the FFT output Z_k (swizzled so step 3's per-lane k ranges hit
    // distinct banks).
    __device__ static constexpr int posZ(int k) {
        return KC > 1 ? ((k & ~(KC - 1)) | ((k ^ (k >> LOGKC)) & (KC - 1))) : k;
    }
    // Slot that receives the transform output X_j.  Pencils are independent, so
    // the outputs may sit in any fixed permutation between passes.  Step 3's
    // lane t writes j = N2*t + c (c fixed per store), so XOR-ing the low bits of
    // j with t makes those stores hit distinct banks, while runs of consecutive
    // j stay conflict-free for the natural-order reads.  sigma is an
    // involution and sigma(0) = 0.
    static constexpr int LOGN2 = ilog2(N2);
    __device__ static constexpr int sigma(int j) { return j ^ ((j >> LOGN2) & (N1 - 1)); }
};

// Per-lane constants of the DST (the lane's index t inside its job is fixed for
// the whole kernel): twiddles w_N^{t k2}, k2 < N2, and shifted sines
// sin(pi (t+N1 j2)/N) + 1/2.
template <typename T, int LOGN>
struct DstLane {
    using S = DstShape<LOGN>;
    T twr[S::N2], twi[S::N2], sp[S::N2];
    // tables (global memory): cos(2 pi m/N), -sin(2 pi m/N), sin(pi m/N) + 1/2, m < N
    __device__ __forceinline__ void init(const T* __restrict__ tables, const int t) {
        constexpr int N = S::N, N1 = S::N1, N2 = S::N2;
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) {
            twr[k2] = tables[t * k2];
            twi[k2] = tables[N + t * k2];
            sp[k2] = tables[2 * N + t + N1 * k2];
        }
    }
    __device__ __forceinline__ T sine(const int j2) const { return sp[j2]; }
};

// The same constants with the shifted sines read from a per-block LDS copy of
// the table at each use instead of held in registers (N2 registers of T fewer
// per lane): for kernels whose register budget is taken by other work.
template <typename T, int LOGN>
struct DstLaneLdsSine {
    using S = DstShape<LOGN>;
    T twr[S::N2], twi[S::N2];
    const T* sl;   // LDS: sin(pi m/N) + 1/2, m < N, from the lane's own m = t
    // lds_sines: N values of the block's LDS, filled by fill_sines before the
    // first barrier
    __device__ __forceinline__ void init(const T* __restrict__ tables, const T* lds_sines, const int t) {
        constexpr int N = S::N, N2 = S::N2;
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) {
            twr[k2] = tables[t * k2];
            twi[k2] = tables[N + t * k2];
        }
        sl = lds_sines + t;
    }
    __device__ static void fill_sines(const T* __restrict__ tables, T* lds_sines, const int tid, const int nt) {
        for (int m = tid; m < S::N; m += nt) lds_sines[m] = tables[2 * S::N + m];
    }
    __device__ __forceinline__ T sine(const int j2) const { return sl[S::N1 * j2]; }
};

// LDS layouts of a set of pencils.  Measured on gfx950: b32 accesses and b64
// reads see 32 banks per half-wavefront (b64 reads in 8-byte banks), b64
// writes 16 eight-byte banks per quarter-wavefront.  A job's N1 lanes touch N1
// slots that are distinct mod N1, so a row pitch congruent to N1 mod 32 makes
// row-pencil (x) accesses conflict-free.  Column-pencil (y) accesses are then
// spread by shifting each row by f(row) < N1 slots (the pitch leaves room).
// The shift is additive rather than an XOR so that, for a fixed row, slot
// offsets stay compile-time constants a ds instruction can carry.
template <int N1>
struct RowSwizzle {
    static constexpr int QJ = 32 / N1;                            // jobs per half-wavefront
    static constexpr int LS = ilog2(QJ);
    static constexpr int FM = (N1 * N1 > 32) ? (N1 * N1) / 32 - 1 : 0;
    static constexpr int FMAX = FM << LS;                         // largest f value
    __device__ static constexpr int f(int row) { return ((row >> LS) & FM) << LS; }
};

// Hides a value's derivation from the optimizer (no instruction is emitted).
__device__ __forceinline__ int opaque(int x) {
    asm("" : "+v"(x));
    return x;
}

// Plane of rows with pitch P; element (row, col) at row*P + col + f(row).
// Pencils are rows (x pass) or columns (y pass).
template <int P, int N1, bool ALONG_X>
struct PlaneLayout {
    using Sw = RowSwizzle<N1>;
    // pencil of job jj, and the distance to its second pencil: jobs in
    // order, pencils jj and jj + JOBS
    __device__ static constexpr int pencil(int jj) { return jj; }
    static constexpr int pair_delta(int jobs) { return jobs; }
    __device__ static constexpr int at(int row, int col) { return row * P + col + Sw::f(row); }
    __device__ static constexpr int addr(int pencil, int slot) {
        return ALONG_X ? at(pencil, slot) : at(slot, pencil);
    }
    // addr(pencil + DP, lv + cp) for a lane-dependent pencil and lv and
    // compile-time DP and cp, cp a multiple of N1.  f() has period N1 (it reads
    // only bits below log2(N1)), so f(lv + cp) = f(lv), and likewise
    // f(pencil + DP) = f(pencil) when N1 divides DP: the constant parts then
    // become an offset the ds instruction carries, and the lane part is
    // computed once per distinct (pencil, lv).  The lane part goes through
    // opaque() so the compiler cannot fold the constants back into the
    // row multiply.
    template <int DP>
    __device__ static int addr_split(int pencil, int lv, int cp) {
        if constexpr (ALONG_X) {
            if constexpr (DP % N1 == 0 || Sw::FM == 0)
                return opaque(pencil * P + Sw::f(pencil) + lv) + (DP * P + cp);
            else
                return opaque((pencil + DP) * P + Sw::f(pencil + DP) + lv) + cp;
        } else {
            return opaque(lv * P + Sw::f(lv) + pencil) + (DP + cp * P);
        }
    }
};

// FP64 plane of N = 128 (N1 = 8, 64 jobs per pass): no row swizzle, a pitch
// P = 2 (mod 32) slots, and jobs dealt to pencils out of order so that both
// the b64 reads (32 eight-byte banks per half-wavefront: 4 jobs) and the b64
// writes (16 per quarter-wavefront: 2 jobs) of either pass are
// conflict-free.  A job's 8 lanes touch slots 8m + (t ^ c) of its pencils
// (m fixed or t-dependent): along a row those are 8 distinct banks, down a
// column rows r with 2r mod 32 spread over 8 distinct even banks.  So
//   x pass (pencils = rows): a job takes rows r and r + 1 (r even: the
//     second pencil one row on, an offset a ds instruction carries); the 4
//     jobs of a half-wavefront take r = e, e+4, e+8, e+12 (2r = 2e + 0, 8,
//     16, 24 mod 32), the 2 of a quarter-wavefront rows 4 apart (8 banks
//     apart mod 16);
//   y pass (pencils = columns): a job takes columns c and c + 64; the 4 jobs
//     of a half-wavefront take c, c+1, c+16, c+17 (odd/even and low/high
//     halves of the 32 banks), the 2 of a quarter-wavefront c, c+1.
// 128 rows of 130 slots: 133 KB for a FP64 plane (the swizzled layout's
// pitch 136 takes 139 KB).
template <int P, bool ALONG_X>
struct PlaneLayoutB64 {
    static_assert(P % 32 == 2, "pitch 2 mod 32");
    __device__ static constexpr int pencil(int jj) {
        const int q = jj >> 2, i = jj & 3;
        return ALONG_X ? 16 * (q >> 1) + 2 * (q & 1) + 4 * i : 32 * (q >> 3) + 2 * (q & 7) + (i & 1) + 16 * (i >> 1);
    }
    static constexpr int pair_delta(int jobs) { return ALONG_X ? 1 : jobs; }
    __device__ static constexpr int at(int row, int col) { return row * P + col; }
    __device__ static constexpr int addr(int pencil, int slot) {
        return ALONG_X ? at(pencil, slot) : at(slot, pencil);
    }
    template <int DP>
    __device__ static int addr_split(int pencil, int lv, int cp) {
        if constexpr (ALONG_X) return opaque(pencil * P + lv) + (DP * P + cp);
        else return opaque(lv * P + pencil) + (DP + cp * P);
    }
};

// v from lane - D within the lane's row of 16 (DPP row_shr:D, a VALU move:
// no LDS instruction as a shuffle through ds_bpermute would take); lanes
// whose source falls outside the row get v back.
template <int D, typename T>
__device__ __forceinline__ T row_shr(T v) {
    static_assert(D >= 1 && D <= 15, "row_shr range");
    if constexpr (sizeof(T) == 8) {
        const long long b = __builtin_bit_cast(long long, v);
        const int lo = __builtin_amdgcn_update_dpp((int)b, (int)b, 0x110 + D, 0xf, 0xf, false);
        const int hi = __builtin_amdgcn_update_dpp((int)(b >> 32), (int)(b >> 32), 0x110 + D, 0xf, 0xf, false);
        return __builtin_bit_cast(T, (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
    } else {
        const int b = __builtin_bit_cast(int, v);
        return __builtin_bit_cast(T, __builtin_amdgcn_update_dpp(b, b, 0x110 + D, 0xf, 0xf, false));
    }
}

// Inclusive scan across `width` consecutive lanes (width a power of two <= 64).
// Within a row of 16 lanes (width <= 16) the lane - d values come by DPP;
// a lane takes one only when lane - d is in its own group (t >= d), so the
// row_shr values from the neighbouring group are never used.
template <int WIDTH, typename T>
__device__ __forceinline__ T lane_scan_incl(T v) {
    if constexpr (WIDTH <= 16) {
        const int t = threadIdx.x & (WIDTH - 1);
        if constexpr (WIDTH > 1) {
            const T o = row_shr<1>(v);
            if (t >= 1) v += o;
        }
        if constexpr (WIDTH > 2) {
            const T o = row_shr<2>(v);
            if (t >= 2) v += o;
        }
        if constexpr (WIDTH > 4) {
            const T o = row_shr<4>(v);
            if (t >= 4) v += o;
        }
        if constexpr (WIDTH > 8) {
            const T o = row_shr<8>(v);
            if (t >= 8) v += o;
        }
    } else {
#pragma unroll
        for (int d = 1; d < WIDTH; d <<= 1) {
            T o = __shfl_up(v, d, WIDTH);
            if ((threadIdx.x & (WIDTH - 1)) >= d) v += o;
        }
    }
    return v;
}

// Orders the LDS accesses of one wavefront across lanes.  A job's lanes lie in
// one wavefront, whose LDS instructions execute in issue order, so only the
// compiler's code motion has to be fenced.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Pencil transform kinds.
//   kDst:     DST-I, FFT<T>::sFFT (reference src/fft.cpp:293-365); N-1 points
//             in slots 1..N-1 (slot 0 is a dummy).
//   kRdftFwd: half-complex real DFT, FFT<T>::pFFT_1 (src/fft.cpp:108-193):
//             X[0] = s*sum x_j, X[k] = s*sum x_j cos(2 pi k j/N),
//             X[N-k] = s*sum x_j sin(2 pi k j/N) (0 < k < N/2),
//             X[N/2] = s*sum (-1)^j x_j; N points in slots 0..N-1.
//   kRdftInv: its inverse series, FFT<T>::pFFT (src/fft.cpp:195-212):
//             x_j = s*(X[0]/2 + X[N/2](-1)^j/2
//                   + sum_{0<k<N/2} X[k] cos(2 pi k j/N) + X[N-k] sin(2 pi k j/N)).
// pFFT(pFFT_1(x)) = (N/2) x, the same round-trip factor as the DST-I.

template <int KIND, int LOGN>
struct PencilIO {
    using S = DstShape<LOGN>;
    static constexpr int N = S::N;
    static constexpr int n = KIND == kDst ? N - 1 : N;   // points per pencil
    static constexpr int off = KIND == kDst ? 1 : 0;     // slot of point 0 on input
    // slot that holds output point m (0 <= m < n) after the transform
    __device__ static constexpr int out(int m) { return KIND == kDst ? S::sigma(m + 1) : S::posZ(m); }
};

// Transform of 2*JOBS pencils in LDS: pencil p, slot j at lds[Lay::addr(p, j)];
// job g transforms pencils g and g + JOBS (packed as the real and imaginary
// parts of one complex FFT of length N).  On entry the points sit in natural
// order from slot PencilIO::off (kDst: slot 0 holds a finite value); on exit
// output point m sits in slot PencilIO::out(m), times `scale` (kDst: slot 0
// holds 0).  All NT threads of the block call this; no block barrier inside
// (jobs never straddle a wavefront).
//
// kRdftFwd: Z = FFT(a + i b); A_k = (Z_k + conj Z_{N-k})/2,
//   B_k = (Z_k - conj Z_{N-k})/(2i), X_a[k] = Re A_k, X_a[N-k] = -Im A_k.
//   Step 3 reads Z_k, Z_{N-k} and writes X[k], X[N-k] into the same two slots.
// kRdftInv: C_k = X[k]/2 - i X[N-k]/2 (Hermitian; C_0, C_{N/2} real),
//   w = C^a + i C^b, x^a + i x^b = sum_k w_k e^{+2 pi i jk/N}
//   = conj(FFT(conj w)); step 2 writes the result directly.
template <int KIND, typename T, int LOGN, int JOBS, typename Lay, int NT, typename Lane = DstLane<T, LOGN>>
__device__ __forceinline__ void pencil_pairs(T* __restrict__ lds, const Lane& L, const T scale,
                                             const int tid) {
    using S = DstShape<LOGN>;
    constexpr int N = S::N, N1 = S::N1, N2 = S::N2, KC = S::KC;
    constexpr int WORK = JOBS * N1;
    static_assert(NT % N1 == 0 && 64 % N1 == 0, "jobs must not straddle wavefronts");
    const int t = tid & (N1 - 1);
#pragma unroll 1
    for (int base = 0; base < WORK; base += NT) {
        const int w = base + tid;
        if (w < WORK) {
            const int g = Lay::pencil(w / N1);
// slot = lv + cp: lv lane-dependent, cp a compile-time multiple of N1 (see
// PlaneLayout::addr_split); pencil gb = g + Lay::pair_delta(JOBS) sits at a constant offset
#define PA(lv, cp) lds[Lay::template addr_split<0>(g, (lv), (cp))]
#define PB(lv, cp) lds[Lay::template addr_split<Lay::pair_delta(JOBS)>(g, (lv), (cp))]

            // ---- step 1: pre-processing + N2-point DFTs over j2 (j = t + N1*j2)
            T yr[N2], yi[N2];
            const int tn0 = (N - t) & (N - 1);   // slot N - j for j = t (0 for j = 0)
#pragma unroll
            for (int j2 = 0; j2 < N2; ++j2) {
                const int j = t + N1 * j2;
                if constexpr (KIND == kRdftFwd) {
                    yr[j2] = PA(t, N1 * j2);
                    yi[j2] = PB(t, N1 * j2);
                } else {
                    // N - j = (N1 - t) + (N - N1 - N1*j2) for j2 > 0
                    const int jl = j2 == 0 ? tn0 : N1 - t, jc = j2 == 0 ? 0 : N - N1 - N1 * j2;
                    const T a0 = PA(t, N1 * j2), a1 = PA(jl, jc);
                    const T b0 = PB(t, N1 * j2), b1 = PB(jl, jc);
                    if constexpr (KIND == kDst) {
                        const T s = L.sine(j2);
                        yr[j2] = fma(s, a0 + a1, -a1);
                        yi[j2] = fma(s, b0 + b1, -b1);
                    } else if (2 * j2 < N2) {   // 0 <= j < N/2
                        const bool edge = j == 0;
                        yr[j2] = edge ? T(0.5) * a0 : T(0.5) * (a0 + b1);
                        yi[j2] = edge ? T(-0.5) * b0 : T(0.5) * (a1 - b0);
                    } else {                              // N/2 <= j < N
                        const bool edge = j == N / 2;
                        yr[j2] = edge ? T(0.5) * a0 : T(0.5) * (a1 - b0);
                        yi[j2] = edge ? T(-0.5) * b0 : T(-0.5) * (a0 + b1);
                    }
                }
            }
            dft<N2>(yr, yi);
#pragma unroll
            for (int k2 = 1; k2 < N2; ++k2) {
                const T c = L.twr[k2], s = L.twi[k2];
                const T r = yr[k2] * c - yi[k2] * s;
                const T i = yr[k2] * s + yi[k2] * c;
                yr[k2] = r;
                yi[k2] = i;
            }
            wave_lds_sync();
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) {   // posY(t, k2) = N1*k2 + (t ^ (k2 mod N1))
                const int lv = t ^ (k2 & (N1 - 1));
                PA(lv, N1 * k2) = yr[k2];
                PB(lv, N1 * k2) = yi[k2];
            }
            wave_lds_sync();

            // ---- step 2: N1-point DFTs over j1 for k2 = t + N1*r
            constexpr int RR = N2 / N1;
            T zr[N2], zi[N2];
#pragma unroll
            for (int r = 0; r < RR; ++r) {
                T vr[N1], vi[N1];
#pragma unroll
                for (int j1 = 0; j1 < N1; ++j1) {   // posY(j1, k2) = N1*t + (j1 ^ t) + N1*N1*r
                    const int lv = N1 * t + (j1 ^ t);
                    vr[j1] = PA(lv, N1 * N1 * r);
                    vi[j1] = PB(lv, N1 * N1 * r);
                }
                dft<N1>(vr, vi);
#pragma unroll
                for (int k1 = 0; k1 < N1; ++k1) {
                    zr[r * N1 + k1] = vr[k1];
                    zi[r * N1 + k1] = vi[k1];
                }
            }
            wave_lds_sync();
#pragma unroll
            for (int r = 0; r < RR; ++r) {
#pragma unroll
                for (int k1 = 0; k1 < N1; ++k1) {
                    // posZ permutes aligned blocks of KC <= N1 slots, so
                    // posZ(t + cz) - cz < N1 for the multiple of N1 cz
                    const int cz = N1 * r + N2 * k1;
                    const int lv = S::posZ(t + cz) - cz;
                    if constexpr (KIND == kRdftInv) {   // x^a = Re Y, x^b = -Im Y
                        PA(lv, cz) = zr[r * N1 + k1] * scale;
                        PB(lv, cz) = -zi[r * N1 + k1] * scale;
                    } else {
                        PA(lv, cz) = zr[r * N1 + k1];
                        PB(lv, cz) = zi[r * N1 + k1];
                    }
                }
            }
            wave_lds_sync();

            if constexpr (KIND == kRdftFwd) {
                // ---- step 3: split the packed spectra; lane t owns k in [t*KC, (t+1)*KC)
                // and their mirrors N-k (k = 0: the Nyquist slot N/2)
#pragma unroll
                for (int i = 0; i < KC; ++i) {
                    const int k = t * KC + i;
                    const bool k0 = k == 0;
                    const int pk = S::posZ(k), pn = S::posZ(k0 ? N / 2 : N - k);
                    const T zkr = PA(pk, 0), zki = PB(pk, 0);
                    const T znr = PA(pn, 0), zni = PB(pn, 0);
                    PA(pk, 0) = (k0 ? zkr : T(0.5) * (zkr + znr)) * scale;
                    PB(pk, 0) = (k0 ? zki : T(0.5) * (zki + zni)) * scale;
                    PA(pn, 0) = (k0 ? znr : T(0.5) * (zni - zki)) * scale;
                    PB(pn, 0) = (k0 ? zni : T(0.5) * (zkr - znr)) * scale;
                }
                wave_lds_sync();
            } else if constexpr (KIND == kDst) {
                // ---- step 3: even outputs directly, odd outputs by prefix sum.
                // The 1/2 of both formulas is applied with the scale at the
                // store (a power of two: the results are bitwise the same).
                const T hs = T(0.5) * scale;
                T ae[KC], ao[KC], be[KC], bo[KC];
                T runA = T(0), runB = T(0);
#pragma unroll
                for (int i = 0; i < KC; ++i) {
                    const int k = t * KC + i;
                    const int pk = S::posZ(k), pn = S::posZ((N - k) & (N - 1));
                    const T zkr = PA(pk, 0), zki = PB(pk, 0);
                    const T znr = PA(pn, 0), zni = PB(pn, 0);
                    ae[i] = zni - zki;
                    be[i] = zkr - znr;
                    T dA = zkr + znr;
                    T dB = zki + zni;
                    if (k == 0) {
                        dA *= T(0.5);
                        dB *= T(0.5);
                    }
                    runA += dA;
                    runB += dB;
                    ao[i] = runA;
                    bo[i] = runB;
                }
                if constexpr (N1 > 1) {
                    const T offA = lane_scan_incl<N1>(runA) - runA;
                    const T offB = lane_scan_incl<N1>(runB) - runB;
#pragma unroll
                    for (int i = 0; i < KC; ++i) {
                        ao[i] += offA;
                        bo[i] += offB;
                    }
                }
                wave_lds_sync();
#pragma unroll
                for (int i = 0; i < KC; ++i) {
                    // sigma(2k + e) = t*N2 + ((2i + e) ^ t) for k = t*KC + i (2i + e < N2, t < N1):
                    // lane part t*N2 + (((2i + e) mod N1) ^ t), the rest a multiple of N1
                    const int k = t * KC + i;
                    const T e_a = k == 0 ? T(0) : ae[i] * hs;
                    const T e_b = k == 0 ? T(0) : be[i] * hs;
                    const int me = 2 * i, mo = 2 * i + 1;
                    const int le = t * N2 + ((me & (N1 - 1)) ^ t), lo = t * N2 + ((mo & (N1 - 1)) ^ t);
                    PA(le, me & ~(N1 - 1)) = e_a;
                    PB(le, me & ~(N1 - 1)) = e_b;
                    PA(lo, mo & ~(N1 - 1)) = ao[i] * hs;
                    PB(lo, mo & ~(N1 - 1)) = bo[i] * hs;
                }
                wave_lds_sync();
            }
#undef PA
#undef PB
        }
    }
}

template <typename T, int LOGN, int JOBS, typename Lay, int NT>
__device__ __forceinline__ void dst_pairs(T* __restrict__ lds, const DstLane<T, LOGN>& L, const T scale,
                                          const int tid) {
    pencil_pairs<kDst, T, LOGN, JOBS, Lay, NT>(lds, L, scale, tid);
}

// The DST of the rows (COLS false: PlaneLayout<P, 8, true>) or the columns
// (COLS true: PlaneLayout<P, 8, false>) of a 128-wide plane (N = 128, N1 = 8,
// N2 = 16, KC = 8), the same transform as pencil_pairs<kDst> with cheaper
// addressing.  Every slot a job's lane t touches is 8m + (t ^ c), c a
// compile-time constant and m either a constant or t, 2t or 15 - t; the row
// swizzle f() has period 8, so the byte offset is (a per-lane base for m) +
// H[c] + (8 slots) * (constant part of m), H[c] = offset of slot t ^ c: one
// add per distinct (base, c) where the generic path computes the slot and
// its swizzle per access.
// B64: the plane is a PlaneLayoutB64 (no row swizzle, jobs dealt to pencils
// by its pencil()) instead of a swizzled PlaneLayout.
template <typename T, int P, int JOBS, int NT, bool COLS, typename Lane = DstLane<T, 7>, bool B64 = false>
__device__ __forceinline__ void dst_pairs128(T* __restrict__ lds, const Lane& L, const T scale,
                                             const int tid) {
    using S = DstShape<7>;
    constexpr int N = 128, N1 = 8, N2 = 16, KC = 8;
    static_assert(S::N1 == N1 && S::N2 == N2 && S::KC == KC, "N = 128 shape");
    static_assert(JOBS * N1 == NT, "one job per thread");
    using Sw = RowSwizzle<N1>;
    constexpr int ES = (int)sizeof(T);
    constexpr int ROW8 = COLS ? 8 * P * ES : 8 * ES;   // bytes of 8 slots
    constexpr int PBO = COLS ? JOBS * ES : JOBS * P * ES;   // pencil g + JOBS (f(g + JOBS) = f(g))
    static_assert(JOBS % N1 == 0, "swizzle period");
    const int t = tid & (N1 - 1);
    const int g = [&] {
        if constexpr (B64) return PlaneLayoutB64<P, !COLS>::pencil(tid / N1);
        else return tid / N1;
    }();
    char* const base = reinterpret_cast<char*>(lds);
    auto fsw = [](int r) { return B64 ? 0 : Sw::f(r); };
    auto rowb = [&](int r) { return COLS ? (r * P + fsw(r)) * ES : r * ES; };   // slot r of pencil 0
    int H[N1];
#pragma unroll
    for (int c = 0; c < N1; ++c) H[c] = opaque(rowb(t ^ c));
    const int pb = COLS ? g * ES : (g * P + fsw(g)) * ES;
    const int b0 = opaque(pb), b1 = opaque(pb + ROW8 * t), b2 = opaque(pb + 2 * ROW8 * t),
              b3 = opaque(pb + ROW8 * (15 - t));
    // the slots outside the 8m + (t ^ c) form: N - t (step 1, j = t) and the
    // mirror of k = 8t (step 3)
    const int rA = opaque(pb + rowb((N - t) & (N - 1)));
    const int rB = opaque(pb + rowb(8 - t));   // f(8) = f(0): row 8(15 - j2) + 8 - t for every t
    const int rC = opaque(pb + rowb(S::posZ((N - 8 * t) & (N - 1))));
    // A ds instruction carries a byte offset below 64 KiB.  When 8 row groups
    // of the plane span more (FP64 columns), the accesses 8..15 row groups on
    // go from bases set 8 groups further (one add each, once per pass) instead
    // of an add per access.
    constexpr bool HI = COLS && 15 * ROW8 > 65535;
    const int b0h = HI ? opaque(pb + 8 * ROW8) : b0, rBh = HI ? opaque(rB + 8 * ROW8) : rB,
              b1h = HI ? opaque(b1 + 8 * ROW8) : b1;
    // lane base + c: c < 8 groups from `lo`, else from `hi` (8 groups on)
    auto split = [&](int lo, int hi, int lane, int groups) {
        return (HI && groups >= 8) ? hi + lane + ROW8 * (groups - 8) : lo + lane + ROW8 * groups;
    };
    auto A = [&](int off) -> T& { return *reinterpret_cast<T*>(base + off); };
    auto B = [&](int off) -> T& { return *reinterpret_cast<T*>(base + off + PBO); };

    // ---- step 1: pre-processing + 16-point DFTs over j2 (j = t + 8 j2)
    T yr[N2], yi[N2];
#pragma unroll
    for (int j2 = 0; j2 < N2; ++j2) {
        const int o0 = split(b0, b0h, H[0], j2);
        const int o1 = j2 == 0 ? rA : split(rB, rBh, 0, 15 - j2);
        const T a0 = A(o0), a1 = A(o1), c0 = B(o0), c1 = B(o1);
        const T sp = L.sine(j2);
        yr[j2] = fma(sp, a0 + a1, -a1);
        yi[j2] = fma(sp, c0 + c1, -c1);
    }
    dft<N2>(yr, yi);
#pragma unroll
    for (int k2 = 1; k2 < N2; ++k2) {
        const T c = L.twr[k2], s = L.twi[k2];
        const T r = yr[k2] * c - yi[k2] * s;
        const T i = yr[k2] * s + yi[k2] * c;
        yr[k2] = r;
        yi[k2] = i;
    }
    wave_lds_sync();
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) {   // slot 8 k2 + (t ^ (k2 mod 8))
        const int o = b0 + H[k2 & 7] + ROW8 * k2;
        A(o) = yr[k2];
        B(o) = yi[k2];
    }
    wave_lds_sync();

    // ---- step 2: 8-point DFTs over j1 for k2 = t + 8 r (slot 8t + (j1 ^ t) + 64 r)
    T zr[N2], zi[N2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        T vr[N1], vi[N1];
#pragma unroll
        for (int j1 = 0; j1 < N1; ++j1) {
            const int o = split(b1, b1h, H[j1], 8 * r);
            vr[j1] = A(o);
            vi[j1] = B(o);
        }
        dft<N1>(vr, vi);
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) {
            zr[r * N1 + k1] = vr[k1];
            zi[r * N1 + k1] = vi[k1];
        }
    }
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < 2; ++r) {
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) {   // posZ(t + cz) = cz + (t ^ ((r + 2 k1) mod 8)), cz = 8 (r + 2 k1)
            const int m = r + 2 * k1;
            const int o = split(b0, b0h, H[m & 7], m);
            A(o) = zr[r * N1 + k1];
            B(o) = zi[r * N1 + k1];
        }
    }
    wave_lds_sync();

    // ---- step 3: even outputs directly, odd outputs by prefix sum (as pencil_pairs)
    const T hs = T(0.5) * scale;
    T ae[KC], ao[KC], be[KC], bo[KC];
    T runA = T(0), runB = T(0);
#pragma unroll
    for (int i = 0; i < KC; ++i) {
        // posZ(8t + i) = 8t + (i ^ t); posZ(N - 8t - i) = 8 (15 - t) + (t ^ (7 ^ (8 - i))) for i > 0
        const int ok = b1 + H[i];
        const int on = i == 0 ? rC : b3 + H[7 ^ (8 - i)];
        const T zkr = A(ok), zki = B(ok);
        const T znr = A(on), zni = B(on);
        ae[i] = zni - zki;
        be[i] = zkr - znr;
        T dA = zkr + znr;
        T dB = zki + zni;
        if (i == 0 && t == 0) {
            dA *= T(0.5);
            dB *= T(0.5);
        }
        runA += dA;
        runB += dB;
        ao[i] = runA;
        bo[i] = runB;
    }
    const T offA = lane_scan_incl<N1>(runA) - runA;
    const T offB = lane_scan_incl<N1>(runB) - runB;
#pragma unroll
    for (int i = 0; i < KC; ++i) {
        ao[i] += offA;
        bo[i] += offB;
    }
    wave_lds_sync();
#pragma unroll
    for (int i = 0; i < KC; ++i) {
        // sigma(2k + e) for k = 8t + i: 16t + ((2i + e) mod 8 ^ t) + 8 ((2i + e) div 8)
        const bool k0 = i == 0 && t == 0;
        const int me = 2 * i, mo = 2 * i + 1;
        const int oe = b2 + H[me & 7] + ROW8 * (me >> 3), oo = b2 + H[mo & 7] + ROW8 * (mo >> 3);
        A(oe) = k0 ? T(0) : ae[i] * hs;
        B(oe) = k0 ? T(0) : be[i] * hs;
        A(oo) = ao[i] * hs;
        B(oo) = bo[i] * hs;
    }
    wave_lds_sync();
}

// ---------------------------------------------------------------------------
// Constant-coefficient cyclic reduction along z for one (kx,ky) mode.
//
// After the DST in x and y, mode (kx,ky) obeys the Toeplitz tridiagonal system
//   b u_{i-1} + d u_i + b u_{i+1} = g_i,  i = 0..n-1,  n = 2^Q - 1,
//   b = 1/hz^2, d = -2/hz^2 - lambda_x[kx] - lambda_y[ky], u_{-1} = u_n = 0,
// the system the reference's z DST + divide + inverse z DST solves
// (src/lapl_cube.cpp:13-30, 70-80, 124-141).  Odd-even cyclic reduction
// (Hockney & Jesshope, the reference's cyclic_reduction,
// src/cyclic_reduction.h:20-56) keeps the coefficients uniform at every level
// when n = 2^Q - 1, so each level is two scalars per mode:
//   alpha_l = -c_{l-1}/d_{l-1},  d_l = d_{l-1} - 2 c_{l-1}^2/d_{l-1},
//   c_l = -c_{l-1}^2/d_{l-1}.
// G lanes share one system: lane c owns the contiguous chunk of rows
// [c*C, (c+1)*C - 1) plus the separator row (c+1)*C - 1 (C = 2^Q/G).  The
// lowest Q-g levels stay inside a chunk; the separators' updates are additive
// so each lane accumulates its contribution to its two neighbouring
// separators and one shuffle combines them; the remaining g-1 levels over the
// G-1 separators are solved redundantly in every lane from shuffled values.
// ---------------------------------------------------------------------------
template <typename T, int Q>
struct CrLevels {
    T alpha[Q > 1 ? Q : 2];   // alpha[l], l = 1..Q-1
    T c[Q];                   // c[l],  l = 0..Q-1
    T inv_d[Q];               // 1/d[l], l = 0..Q-1
};

template <typename T, int Q>
__device__ __forceinline__ void cr_levels(CrLevels<T, Q>& L, T b, T d) {
    T cl = b, dl = d;
    L.c[0] = cl;
    L.inv_d[0] = T(1) / dl;
#pragma unroll
    for (int l = 1; l < Q; ++l) {
        const T t = cl * cl * L.inv_d[l - 1];
        L.alpha[l] = -cl * L.inv_d[l - 1];
        dl = dl - T(2) * t;
        cl = -t;
        L.c[l] = cl;
        L.inv_d[l] = T(1) / dl;
    }
}

}  // namespace fdmgpu
