// CPU oracle for the fdm fast Poisson solve -- TEST INFRASTRUCTURE ONLY.
//
// This file is a from-scratch CPU restatement of the reference algorithms on
// the LaplCube / LaplRect solve() path.  It is the checker the GPU library is
// compared against; it is never linked into, loaded by, or called from the
// product library (fdm_amd/csrc).  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it.
//
// Every function cites the reference file:line it follows.  Parity of each
// function is pinned by tests/test_oracle.py (see the PINNED-BY line above
// each entry point):
//   * against oracle/_ref (the reference's own sources compiled by
//     oracle/Makefile) where that builds,
//   * against committed golden vectors in tests/golden/,
//   * against closed forms / library routines (scipy.fft.dst, scipy dgtsv,
//     dense numpy solves, manufactured solutions).
//
// The 1-D sine/cosine/real-DFT transforms here are computed with an
// independent algorithm (odd/even extension + radix-2 complex FFT), not the
// reference's Samarskii-Nikolaev recursion (src/fft.cpp:293-445), so results
// agree with the reference to rounding, not bitwise.  The solver pass order,
// scale factors, eigenvalue tables, index quirks and tridiagonal elimination
// orders follow the reference exactly.
//
// Build: oracle/Makefile (g++ -O3 -fopenmp -shared).

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------
// Radix-2 complex FFT (forward, e^{-2 pi i jk/L}), L = 2^m, in T arithmetic.
// Twiddles computed in double and rounded to T.
// ---------------------------------------------------------------------------
template <typename T>
struct CFFT {
    int L = 0;
    std::vector<std::complex<T>> w;   // w[k] = exp(-2 pi i k / L), k < L/2
    std::vector<int> rev;

    explicit CFFT(int L_) : L(L_), w(std::max(1, L_ / 2)), rev(L_) {
        for (int k = 0; k < L / 2; k++) {
            double a = -2.0 * M_PI * k / L;
            w[k] = std::complex<T>((T)std::cos(a), (T)std::sin(a));
        }
        int m = 0;
        while ((1 << m) < L) m++;
        for (int i = 0; i < L; i++) {
            int r = 0;
            for (int b = 0; b < m; b++) if (i & (1 << b)) r |= 1 << (m - 1 - b);
            rev[i] = r;
        }
    }

    void run(std::complex<T>* a) const {
        for (int i = 0; i < L; i++) if (i < rev[i]) std::swap(a[i], a[rev[i]]);
        for (int len = 2; len <= L; len <<= 1) {
            int half = len >> 1, step = L / len;
            for (int i = 0; i < L; i += len) {
                for (int k = 0; k < half; k++) {
                    std::complex<T> t = a[i + k + half] * w[k * step];
                    std::complex<T> u = a[i + k];
                    a[i + k] = u + t;
                    a[i + k + half] = u - t;
                }
            }
        }
    }
};

bool is_pow2(int N) { return N >= 1 && (N & (N - 1)) == 0; }

// ---------------------------------------------------------------------------
// DST-I of two pencils at once (follows the definition of FFT<T>::sFFT,
// src/fft.cpp:293-365 / sFT src/asp_fft.cpp:308-318):
//   X[k] = scale * sum_{j=1}^{N-1} x[j] sin(pi j k / N),  k = 1..N-1.
// Inputs/outputs indexed 1..N-1 (element 0 untouched), like the reference.
// Method: odd extension to length 2N, packed as (xa + i xb), one complex FFT:
//   FFT(odd(xa) + i odd(xb))[k] = -2i DSTa[k] + 2 DSTb[k].
// ---------------------------------------------------------------------------
template <typename T>
void dst_pair(const CFFT<T>& f2n, int N, const T* xa, const T* xb, T* Xa, T* Xb,
              T scale, std::complex<T>* buf) {
    const int L = 2 * N;
    buf[0] = 0;
    buf[N] = 0;
    for (int j = 1; j < N; j++) {
        T a = xa[j], b = xb ? xb[j] : T(0);
        buf[j] = std::complex<T>(a, b);
        buf[L - j] = std::complex<T>(-a, -b);
    }
    f2n.run(buf);
    const T h = scale * T(0.5);
    for (int k = 1; k < N; k++) {
        Xa[k] = -buf[k].imag() * h;
        if (Xb) Xb[k] = buf[k].real() * h;
    }
}

// DCT-I (follows FFT<T>::cFFT, src/fft.cpp:367-445 / cFT src/asp_fft.cpp:404-416):
//   X[k] = scale*(0.5 x0 + sum_{j=1}^{N-1} x_j cos(pi jk/N) + 0.5 (-1)^k x_N), k=0..N.
template <typename T>
void dct1(const CFFT<T>& f2n, int N, const T* x, T* X, T scale, std::complex<T>* buf) {
    const int L = 2 * N;
    buf[0] = x[0];
    buf[N] = x[N];
    for (int j = 1; j < N; j++) {
        buf[j] = x[j];
        buf[L - j] = x[j];
    }
    f2n.run(buf);
    for (int k = 0; k <= N; k++) X[k] = buf[k].real() * scale * T(0.5);
}

// Periodic forward real DFT in half-complex layout (follows FFT<T>::pFFT_1,
// src/fft.cpp:108-193):
//   X[0] = s*sum x_j, X[k] = s*sum x_j cos(2pi jk/N), X[N-k] = s*sum x_j sin(2pi jk/N),
//   X[N/2] = s*sum (-1)^j x_j.
template <typename T>
void rdft_fwd(const CFFT<T>& fn, int N, const T* x, T* X, T scale, std::complex<T>* buf) {
    for (int j = 0; j < N; j++) buf[j] = x[j];
    fn.run(buf);
    X[0] = buf[0].real() * scale;
    if (N == 1) return;
    X[N / 2] = buf[N / 2].real() * scale;
    for (int k = 1; k < N / 2; k++) {
        X[k] = buf[k].real() * scale;
        X[N - k] = -buf[k].imag() * scale;
    }
}

// Periodic inverse series (follows FFT<T>::pFFT, src/fft.cpp:195-212):
//   x_j = s*(0.5 X0 + sum_{k=1}^{N/2-1} (X_k cos(2pi jk/N) + X_{N-k} sin(2pi jk/N))
//            + 0.5 (-1)^j X_{N/2}),
// so that pFFT(pFFT_1(x, 1), 2/N) = x (ut/ut_fft.cpp:73-77).
template <typename T>
void rdft_inv(const CFFT<T>& fn, int N, const T* X, T* x, T scale, std::complex<T>* buf) {
    if (N == 1) { x[0] = X[0] * scale * T(0.5); return; }
    // x_j = Re sum_k c_k e^{+2 pi i jk/N} with c_0 = X0/2, c_{N/2} = X_{N/2}/2,
    // c_k = (X_k - i X_{N-k})/2 and c_{N-k} = conj(c_k) for 0 < k < N/2.
    // Forward FFT of conj(c) gives conj of the inverse sum.
    buf[0] = X[0] * T(0.5);
    buf[N / 2] = X[N / 2] * T(0.5);
    for (int k = 1; k < N / 2; k++) {
        std::complex<T> c(X[k] * T(0.5), -X[N - k] * T(0.5));
        buf[k] = std::conj(c);
        buf[N - k] = c;
    }
    fn.run(buf);
    for (int j = 0; j < N; j++) x[j] = buf[j].real() * scale;
}

// Which 1-D transform a LaplCube axis uses.
enum AxisKind { AX_DIRICHLET = 0, AX_PERIODIC = 1 };

// Tensor index wrap of the reference (src/tensor.h:118-128): a periodic axis
// with offsets [lo, hi] maps y -> (y - lo + len) % len + lo.  LaplCube always
// views ans/rhs with offsets {1,n} (src/lapl_cube.h:80) while its loops run
// 0..n-1 on periodic axes, so loop index i addresses array slot
// ((i - 1 + n) % n) (a one-element cyclic shift).
inline int wrap_slot(int i, int n, bool periodic) {
    if (!periodic) return i - 1;            // offsets {1,n}: slot = i - 1
    return ((i - 1) % n + n) % n;
}

// ---------------------------------------------------------------------------
// LaplCube<T>::solve (src/lapl_cube.cpp:9-142) with init_lm
// (src/lapl_cube.cpp:145-172) and the constructor's index bounds / table
// sharing (src/lapl_cube.h:58-100), including its quirks:
//   * x bounds use the z flag (src/lapl_cube.h:68,72,76);
//   * lm_x aliases lm_y when xpoints == ypoints, lm_z aliases lm_y when
//     zpoints == ypoints (src/lapl_cube.cpp:162,171);
//   * zero mode := 0 when every axis is periodic (src/lapl_cube.cpp:82-84).
// Uniform flags only (the only instantiations, src/lapl_cube.cpp:174-182).
// ---------------------------------------------------------------------------
template <typename T>
int lapl_cube_solve(T* ans, const T* rhs, int periodic, double dx, double dy, double dz,
                    double lx, double ly, double lz, int nx, int ny, int nz) {
    const bool per = periodic != 0;
    const int zpoints = per ? nz : nz + 1, ypoints = per ? ny : ny + 1, xpoints = per ? nx : nx + 1;
    if (!is_pow2(zpoints) || !is_pow2(ypoints) || !is_pow2(xpoints)) return 1;
    const int lo = per ? 0 : 1;
    const int zn = per ? nz - 1 : nz, yn = per ? ny - 1 : ny, xn = per ? nx - 1 : nx;
    const double slx = std::sqrt(2. / lx), sly = std::sqrt(2. / ly), slz = std::sqrt(2. / lz);

    // init_lm
    auto lam = [&](int k, int n, double h) -> T {
        double s = per ? std::sin(k * M_PI / n) : std::sin(k * M_PI * 0.5 / (n + 1));
        return (T)(4. / (h * h) * (s * s));
    };
    std::vector<T> lm_y(ny + 1), lm_x_(nx + 1), lm_z_(nz + 1);
    for (int k = lo; k <= yn; k++) lm_y[k] = lam(k, ny, dy);
    for (int j = lo; j <= xn; j++) lm_x_[j] = lam(j, nx, dx);
    for (int i = lo; i <= zn; i++) lm_z_[i] = lam(i, nz, dz);
    const T* lm_x = xpoints == ypoints ? lm_y.data() : lm_x_.data();
    const T* lm_z = zpoints == ypoints ? lm_y.data() : lm_z_.data();

    const size_t total = (size_t)nx * ny * nz;
    std::vector<T> W(total);   // RHSm
    auto at = [&](int i, int k, int j) -> size_t {
        return ((size_t)wrap_slot(i, nz, per) * ny + wrap_slot(k, ny, per)) * nx + wrap_slot(j, nx, per);
    };

    const int maxp = std::max({xpoints, ypoints, zpoints});
    CFFT<T> fx(per ? xpoints : 2 * xpoints), fy(per ? ypoints : 2 * ypoints), fz(per ? zpoints : 2 * zpoints);

    // One 1-D pass along an axis.  src holds pencils addressed by get(), result
    // written through put().  Dirichlet pencils are paired (two per FFT).
    auto pass = [&](const CFFT<T>& f, int points, int outer_lo, int outer_hi, int mid_lo, int mid_hi,
                    int in_lo, int in_hi, T factor, bool inverse, auto index, const T* src, T* dst) {
        const int cnt = (outer_hi - outer_lo + 1) * (mid_hi - mid_lo + 1);
#pragma omp parallel
        {
            std::vector<T> sa(maxp + 1), sb(maxp + 1), Sa(maxp + 1), Sb(maxp + 1);
            std::vector<std::complex<T>> buf(2 * maxp + 2);
#pragma omp for schedule(static)
            for (int pair = 0; pair < (cnt + 1) / 2; pair++) {
                int p0 = 2 * pair, p1 = 2 * pair + 1;
                bool two = p1 < cnt;
                int o0 = outer_lo + p0 / (mid_hi - mid_lo + 1), m0 = mid_lo + p0 % (mid_hi - mid_lo + 1);
                int o1 = two ? outer_lo + p1 / (mid_hi - mid_lo + 1) : 0;
                int m1 = two ? mid_lo + p1 % (mid_hi - mid_lo + 1) : 0;
                for (int t = in_lo; t <= in_hi; t++) {
                    sa[t] = src[index(o0, m0, t)];
                    if (two) sb[t] = src[index(o1, m1, t)];
                }
                if (!per) {
                    dst_pair<T>(f, points, sa.data(), two ? sb.data() : nullptr, Sa.data(),
                                two ? Sb.data() : nullptr, factor, buf.data());
                } else if (!inverse) {
                    rdft_fwd<T>(f, points, sa.data(), Sa.data(), factor, buf.data());
                    if (two) rdft_fwd<T>(f, points, sb.data(), Sb.data(), factor, buf.data());
                } else {
                    rdft_inv<T>(f, points, sa.data(), Sa.data(), factor, buf.data());
                    if (two) rdft_inv<T>(f, points, sb.data(), Sb.data(), factor, buf.data());
                }
                for (int t = in_lo; t <= in_hi; t++) {
                    dst[index(o0, m0, t)] = Sa[t];
                    if (two) dst[index(o1, m1, t)] = Sb[t];
                }
            }
        }
    };

    // pass 1: z forward (src/lapl_cube.cpp:13-30): outer k (y), mid j (x), inner i (z)
    pass(fz, zpoints, lo, yn, lo, xn, lo, zn, (T)(dz * slz), false,
         [&](int k, int j, int i) { return at(i, k, j); }, rhs, W.data());
    // pass 2: y forward (:32-49)
    pass(fy, ypoints, lo, zn, lo, xn, lo, yn, (T)(dy * sly), false,
         [&](int i, int j, int k) { return at(i, k, j); }, W.data(), W.data());
    // pass 3: x forward (:51-68)
    pass(fx, xpoints, lo, zn, lo, yn, lo, xn, (T)(dx * slx), false,
         [&](int i, int k, int j) { return at(i, k, j); }, W.data(), W.data());
    // pass 4: divide (:70-80)
#pragma omp parallel for schedule(static)
    for (int i = lo; i <= zn; i++)
        for (int k = lo; k <= yn; k++)
            for (int j = lo; j <= xn; j++) {
                T k2 = lm_z[i] + lm_y[k] + lm_x[j];
                W[at(i, k, j)] /= -k2;
            }
    if (per) W[at(0, 0, 0)] = 0;   // (:82-84)
    // passes 5-7: inverse x, y, z (:86-141)
    pass(fx, xpoints, lo, zn, lo, yn, lo, xn, (T)slx, true,
         [&](int i, int k, int j) { return at(i, k, j); }, W.data(), ans);
    pass(fy, ypoints, lo, zn, lo, xn, lo, yn, (T)sly, true,
         [&](int i, int j, int k) { return at(i, k, j); }, ans, ans);
    pass(fz, zpoints, lo, yn, lo, xn, lo, zn, (T)slz, true,
         [&](int k, int j, int i) { return at(i, k, j); }, ans, ans);
    return 0;
}

// ---------------------------------------------------------------------------
// LAPACK ?gtsv, NRHS = 1 (called from src/lapl_rect.cpp:90 via src/blas.h:67-75).
// Third-party routine (reference LAPACK dgtsv/sgtsv as shipped in OpenBLAS,
// version unpinned by the reference's CI, .github/workflows/c-cpp.yml:25);
// this restates its published algorithm: Gaussian elimination with partial
// pivoting (row interchange when |d_i| < |dl_i|), then back substitution with
// the second superdiagonal fill.  Returns INFO (0 ok, i>0: U(i,i) is zero).
// PINNED-BY: tests/test_oracle.py::test_gtsv_vs_scipy (scipy.linalg.lapack).
// ---------------------------------------------------------------------------
template <typename T>
int gtsv(int n, T* dl, T* d, T* du, T* b) {
    if (n <= 0) return 0;
    for (int i = 0; i < n - 1; i++) {
        if (std::abs(d[i]) >= std::abs(dl[i])) {
            if (d[i] == T(0)) return i + 1;
            T fact = dl[i] / d[i];
            d[i + 1] -= fact * du[i];
            b[i + 1] -= fact * b[i];
            if (i < n - 2) dl[i] = 0;
        } else {
            T fact = d[i] / dl[i];
            d[i] = dl[i];
            T temp = d[i + 1];
            d[i + 1] = du[i] - fact * temp;
            if (i < n - 2) {
                dl[i] = du[i + 1];
                du[i + 1] = -fact * dl[i];
            }
            du[i] = temp;
            temp = b[i];
            b[i] = b[i + 1];
            b[i + 1] = temp - fact * b[i + 1];
        }
    }
    if (d[n - 1] == T(0)) return n;
    b[n - 1] /= d[n - 1];
    if (n > 1) b[n - 2] = (b[n - 2] - du[n - 2] * b[n - 1]) / d[n - 2];
    for (int i = n - 3; i >= 0; i--) b[i] = (b[i] - du[i] * b[i + 1] - dl[i] * b[i + 2]) / d[i];
    return 0;
}

// ---------------------------------------------------------------------------
// LaplRect<T>::solve (src/lapl_rect.cpp:63-110) with init_lm (:13-22) and
// init_Mat (:44-60); y-periodic when yper != 0.  Layout [ny'][nx], x fastest,
// ny' = ny (Dirichlet) or ny (periodic, rows 0..ny-1).
// lm_y_scale/L_scale/U_scale: length nx+1 arrays (index 0 unused), may be NULL
// (= all ones, the constructor default src/lapl_rect.h:57-59).
// Returns the first non-zero gtsv INFO (the reference aborts on it, :91).
// ---------------------------------------------------------------------------
template <typename T>
int lapl_rect_solve(T* ans, const T* rhs, int yper, double dx, double dy, double lx, double ly,
                    int nx, int ny, const T* lm_y_scale, const T* L_scale, const T* U_scale) {
    const bool per = yper != 0;
    const int y1 = per ? 0 : 1, yn = per ? ny - 1 : ny, ypoints = per ? ny : ny + 1;
    if (!is_pow2(ypoints) || nx < 1) return -1;
    const double dx2 = dx * dx, dy2 = dy * dy;
    const double sly = std::sqrt(2. / ly);
    (void)lx;
    std::vector<T> lm_y(ny + 1);
    for (int k = y1; k <= yn; k++) {
        double s = per ? std::sin(k * M_PI / ny) : std::sin(k * M_PI * 0.5 / (ny + 1));
        lm_y[k] = (T)(4. / dy2 * (s * s));
    }
    std::vector<T> ones(nx + 1, T(1));
    if (!lm_y_scale) lm_y_scale = ones.data();
    if (!L_scale) L_scale = ones.data();
    if (!U_scale) U_scale = ones.data();

    const int rows = yn - y1 + 1;
    std::vector<T> W((size_t)rows * nx);
    auto at = [&](int k, int j) { return (size_t)(k - y1) * nx + (j - 1); };
    CFFT<T> fy(per ? ypoints : 2 * ypoints);

    auto ypass = [&](const T* src, T* dst, T factor, bool inverse) {
#pragma omp parallel
        {
            std::vector<T> sa(ypoints + 1), sb(ypoints + 1), Sa(ypoints + 1), Sb(ypoints + 1);
            std::vector<std::complex<T>> buf(2 * ypoints + 2);
#pragma omp for schedule(static)
            for (int pair = 0; pair < (nx + 1) / 2; pair++) {
                int j0 = 1 + 2 * pair, j1 = j0 + 1;
                bool two = j1 <= nx;
                for (int k = y1; k <= yn; k++) {
                    sa[k] = src[at(k, j0)];
                    if (two) sb[k] = src[at(k, j1)];
                }
                if (!per) {
                    dst_pair<T>(fy, ypoints, sa.data(), two ? sb.data() : nullptr, Sa.data(),
                                two ? Sb.data() : nullptr, factor, buf.data());
                } else if (!inverse) {
                    rdft_fwd<T>(fy, ypoints, sa.data(), Sa.data(), factor, buf.data());
                    if (two) rdft_fwd<T>(fy, ypoints, sb.data(), Sb.data(), factor, buf.data());
                } else {
                    rdft_inv<T>(fy, ypoints, sa.data(), Sa.data(), factor, buf.data());
                    if (two) rdft_inv<T>(fy, ypoints, sb.data(), Sb.data(), factor, buf.data());
                }
                for (int k = y1; k <= yn; k++) {
                    dst[at(k, j0)] = Sa[k];
                    if (two) dst[at(k, j1)] = Sb[k];
                }
            }
        }
    };

    ypass(rhs, W.data(), (T)(dy * sly), false);
    int info_all = 0;
#pragma omp parallel
    {
        std::vector<T> L(nx), D(nx), U(nx);
#pragma omp for schedule(static)
        for (int k = y1; k <= yn; k++) {
            // init_Mat(k)
            for (int j = 1; j <= nx; j++) {
                D[j - 1] = (T)(-2 / dx2 - lm_y[k] * lm_y_scale[j]);
                if (j > 1) L[j - 2] = (T)(L_scale[j] / dx2);
                if (j < nx) U[j - 1] = (T)(U_scale[j] / dx2);
            }
            int info = gtsv<T>(nx, L.data(), D.data(), U.data(), &W[at(k, 1)]);
            if (info != 0) {
#pragma omp critical
                if (info_all == 0) info_all = info;
            }
        }
    }
    if (info_all) return info_all;
    ypass(W.data(), ans, (T)sly, true);
    return 0;
}

// ---------------------------------------------------------------------------
// LaplRectFFT2<T>::solve (src/lapl_rect.cpp:113-206) with its init_lm
// (:25-41).  Flags: yper (F::head), xper (F::tail::head).
// ---------------------------------------------------------------------------
template <typename T>
int lapl_rect_fft2_solve(T* ans, const T* rhs, int yper, int xper, double dx, double dy, double lx,
                         double ly, int nx, int ny, const T* lm_y_scale) {
    const int y1 = yper ? 0 : 1, yn = yper ? ny - 1 : ny, ypoints = yper ? ny : ny + 1;
    const int x1 = xper ? 0 : 1, xn = xper ? nx - 1 : nx, xpoints = xper ? nx : nx + 1;
    if (!is_pow2(ypoints) || !is_pow2(xpoints)) return -1;
    const double dx2 = dx * dx, dy2 = dy * dy;
    const double slx = std::sqrt(2. / lx), sly = std::sqrt(2. / ly);
    std::vector<T> lm_y(ny + 1), lm_x_(xpoints);
    for (int k = y1; k <= yn; k++) {
        double s = yper ? std::sin(k * M_PI / ny) : std::sin(k * M_PI * 0.5 / (ny + 1));
        lm_y[k] = (T)(4. / dy2 * (s * s));
    }
    for (int j = x1; j <= xn; j++) {
        double s = xper ? std::sin(j * M_PI / xpoints) : std::sin(j * M_PI * 0.5 / xpoints);
        lm_x_[j] = (T)(4. / dx2 * (s * s));
    }
    // lm_x aliases lm_y only for the all-Dirichlet instantiation (:36-40)
    const T* lm_x = (!yper && !xper && xpoints == ypoints) ? lm_y.data() : lm_x_.data();
    std::vector<T> ones(nx + 1, T(1));
    if (!lm_y_scale) lm_y_scale = ones.data();

    const int rows = yn - y1 + 1, cols = xn - x1 + 1;
    std::vector<T> W((size_t)rows * cols);
    auto at = [&](int k, int j) { return (size_t)(k - y1) * cols + (j - x1); };
    const int maxp = std::max(xpoints, ypoints);

    auto pass1d = [&](bool alongy, bool perax, int points, const T* src, T* dst, T factor, bool inverse) {
        CFFT<T> f(perax ? points : 2 * points);
        const int cnt = alongy ? cols : rows;
#pragma omp parallel
        {
            std::vector<T> sa(maxp + 1), Sa(maxp + 1);
            std::vector<std::complex<T>> buf(2 * maxp + 2);
#pragma omp for schedule(static)
            for (int p = 0; p < cnt; p++) {
                int lo = alongy ? y1 : x1, hi = alongy ? yn : xn;
                for (int t = lo; t <= hi; t++) sa[t] = src[alongy ? at(t, x1 + p) : at(y1 + p, t)];
                if (!perax) dst_pair<T>(f, points, sa.data(), nullptr, Sa.data(), nullptr, factor, buf.data());
                else if (!inverse) rdft_fwd<T>(f, points, sa.data(), Sa.data(), factor, buf.data());
                else rdft_inv<T>(f, points, sa.data(), Sa.data(), factor, buf.data());
                for (int t = lo; t <= hi; t++) dst[alongy ? at(t, x1 + p) : at(y1 + p, t)] = Sa[t];
            }
        }
    };
    pass1d(true, yper, ypoints, rhs, W.data(), (T)(dy * sly), false);
    pass1d(false, xper, xpoints, W.data(), W.data(), (T)(dx * slx), false);
    for (int k = y1; k <= yn; k++)
        for (int j = x1; j <= xn; j++) W[at(k, j)] /= -lm_y[k] * lm_y_scale[j] - lm_x[j];
    if (y1 == 0 && x1 == 0) W[at(y1, x1)] = 1;   // (:171)
    pass1d(false, xper, xpoints, W.data(), ans, (T)slx, true);
    pass1d(true, yper, ypoints, ans, ans, (T)sly, true);
    return 0;
}

// ---------------------------------------------------------------------------
// Thomas algorithm solve_tdiag_linear_my / ..f_my (src/asp_gauss.cpp:185-229):
// A1 = sub-diagonal (A1[j] couples row j+1 to x_j), A2 = diagonal,
// A3 = super-diagonal; B overwritten by the solution; A1, A2 clobbered.
// ---------------------------------------------------------------------------
template <typename T>
void thomas(T* B, T* A1, T* A2, T* A3, int n) {
    for (int j = 0; j < n - 1; j++) {
        T a = A1[j] / A2[j];
        A1[j] = 0;
        A2[j + 1] -= A3[j] * a;
        B[j + 1] -= B[j] * a;
    }
    for (int j = n - 1; j > 0; j--) {
        T a = A3[j - 1] / A2[j];
        B[j - 1] -= B[j] * a;
    }
    for (int i = 0; i < n; i++) B[i] *= 1 / A2[i];
}

// ---------------------------------------------------------------------------
// Odd-even cyclic reduction, Hockney & Jesshope form (cyclic_reduction,
// src/cyclic_reduction.h:20-56) and the general-n form
// (cyclic_reduction_general, :59-107).  d = diagonal, e = lower (e[0] = 0),
// f = upper, b = rhs -> solution; q = ceil(log2(n+1)).
// ---------------------------------------------------------------------------
template <typename T>
void cyclic_reduction(T* d, T* e, T* f, T* b, int q, int n, bool general) {
    for (int l = 1; l < q; l++) {
        const int s = 1 << l, h = 1 << (l - 1);
        int j = s - 1;
        for (; j + h < n; j += s) {
            T alpha = -e[j] / d[j - h];
            T gamma = -f[j] / d[j + h];
            d[j] += alpha * f[j - h] + gamma * e[j + h];
            b[j] += alpha * b[j - h] + gamma * b[j + h];
            e[j] = alpha * e[j - h];
            f[j] = gamma * f[j + h];
        }
        if (general) {
            for (; j < n; j += s) {
                T alpha = -e[j] / d[j - h];
                d[j] += alpha * f[j - h];
                b[j] += alpha * b[j - h];
                e[j] = alpha * e[j - h];
            }
        }
    }
    int top = (1 << (q - 1)) - 1;
    if (general) top = std::min(top, n - 1);
    b[top] = b[top] / d[top];
    for (int l = q - 1; l > 0; l--) {
        const int s = 1 << l, h = 1 << (l - 1);
        int j = h - 1;
        // first row of the level has no left neighbour
        for (; j - h < 0 && j < n; j += s) b[j] = (b[j] - f[j] * b[j + h]) / d[j];
        for (; j + h < n; j += s) b[j] = (b[j] - e[j] * b[j - h] - f[j] * b[j + h]) / d[j];
        for (; j < n; j += s) b[j] = (b[j] - e[j] * b[j - h]) / d[j];
    }
}

// ---------------------------------------------------------------------------
// NSCube::poisson RHS assembly (src/ns_cube.cpp:204-235).  Arrays use the
// NSCube tensor offsets (src/ns_cube.h:66-75), x fastest:
//   F {1..nz, 1..ny, 0..nx}, G {1..nz, 0..ny, 1..nx}, H {0..nz, 1..ny, 1..nx},
//   p {0..nz+1, 0..ny+1, 0..nx+1}, RHS {1..nz, 1..ny, 1..nx}.
// ---------------------------------------------------------------------------
template <typename T>
void ns_poisson_rhs(T* RHS, const T* F, const T* G, const T* H, const T* p, int nx, int ny, int nz,
                    double dx, double dy, double dz, double dt) {
    const double dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
    auto Fi = [&](int i, int k, int j) { return F[((size_t)(i - 1) * ny + (k - 1)) * (nx + 1) + j]; };
    auto Gi = [&](int i, int k, int j) { return G[((size_t)(i - 1) * (ny + 1) + k) * nx + (j - 1)]; };
    auto Hi = [&](int i, int k, int j) { return H[((size_t)i * ny + (k - 1)) * nx + (j - 1)]; };
    auto Pi = [&](int i, int k, int j) { return p[((size_t)i * (ny + 2) + k) * (nx + 2) + j]; };
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 1; i <= nz; i++)
        for (int k = 1; k <= ny; k++)
            for (int j = 1; j <= nx; j++) {
                T r = ((Fi(i, k, j) - Fi(i, k, j - 1)) / dx + (Gi(i, k, j) - Gi(i, k - 1, j)) / dy +
                       (Hi(i, k, j) - Hi(i - 1, k, j)) / dz) / dt;
                if (i <= 1) r -= Pi(i - 1, k, j) / dz2;
                if (k <= 1) r -= Pi(i, k - 1, j) / dy2;
                if (j <= 1) r -= Pi(i, k, j - 1) / dx2;
                if (j >= nx) r -= Pi(i, k, j + 1) / dx2;
                if (k >= ny) r -= Pi(i, k + 1, j) / dy2;
                if (i >= nz) r -= Pi(i + 1, k, j) / dz2;
                RHS[((size_t)(i - 1) * ny + (k - 1)) * nx + (j - 1)] = r;
            }
}

// ---------------------------------------------------------------------------
// NSCube<double> time step (src/ns_cube.h:12-78, src/ns_cube.cpp:27-277):
// lid-driven cavity, staggered MAC grid, explicit advection + diffusion
// (F, G, H), pressure Poisson projection through LaplCube, velocity update.
// Arrays (x fastest), with the reference's index ranges:
//   u  i in [0,nz+1], k in [0,ny+1], j in [-1,nx+1]   (src/ns_cube.h:66)
//   v  i in [0,nz+1], k in [-1,ny+1], j in [0,nx+1]   (:67)
//   w  i in [-1,nz+1], k in [0,ny+1], j in [0,nx+1]   (:68)
//   p  i,k,j in [0,n+1]; x, RHS in [1,n]; F j in [0,nx]; G k in [0,ny];
//   H i in [0,nz] (:69-75)
// Quirk kept: the lid loop's x bound is nz+1 (src/ns_cube.cpp:68); a cube
// with nz > nx would index u out of range (the reference's checked tensors
// abort), so such sizes are rejected.
// ---------------------------------------------------------------------------
struct NsState {
    int nx, ny, nz;
    double dx, dy, dz, U0, Re, dt, lx, ly, lz;
    std::vector<double> u, v, w, p, x, F, G, H, RHS;
    long steps = 0;
    double& U(int i, int k, int j) { return u[((size_t)i * (ny + 2) + k) * (nx + 3) + (j + 1)]; }
    double& V(int i, int k, int j) { return v[((size_t)i * (ny + 3) + (k + 1)) * (nx + 2) + j]; }
    double& W(int i, int k, int j) { return w[((size_t)(i + 1) * (ny + 2) + k) * (nx + 2) + j]; }
    double& P(int i, int k, int j) { return p[((size_t)i * (ny + 2) + k) * (nx + 2) + j]; }
    double& X(int i, int k, int j) { return x[((size_t)(i - 1) * ny + (k - 1)) * nx + (j - 1)]; }
    double& Fv(int i, int k, int j) { return F[((size_t)(i - 1) * ny + (k - 1)) * (nx + 1) + j]; }
    double& Gv(int i, int k, int j) { return G[((size_t)(i - 1) * (ny + 1) + k) * nx + (j - 1)]; }
    double& Hv(int i, int k, int j) { return H[((size_t)i * ny + (k - 1)) * nx + (j - 1)]; }
};

static inline double sqd(double a) { return a * a; }

// init_bound (src/ns_cube.cpp:64-127)
void ns_init_bound(NsState& S) {
    const int nx = S.nx, ny = S.ny, nz = S.nz;
    for (int k = 0; k <= ny + 1; k++)
        for (int j = -1; j <= nz + 1; j++) S.U(nz + 1, k, j) = 2 * S.U0 - S.U(nz, k, j);
    for (int i = 0; i <= nz + 1; i++)
        for (int k = 0; k <= ny + 1; k++) {
            S.U(i, k, -1) = S.U(i, k, 1);
            S.U(i, k, nx + 1) = S.U(i, k, nx - 1);
        }
    for (int i = 0; i <= nz + 1; i++)
        for (int j = 0; j <= nx + 1; j++) {
            S.V(i, -1, j) = S.V(i, 1, j);
            S.V(i, ny + 1, j) = S.V(i, ny - 1, j);
        }
    for (int k = 0; k <= ny + 1; k++)
        for (int j = 0; j <= nx + 1; j++) {
            S.W(-1, k, j) = S.W(1, k, j);
            S.W(nz + 1, k, j) = S.W(nz - 1, k, j);
        }
    const double Re = S.Re;
    for (int i = 1; i <= nz; i++)
        for (int k = 1; k <= ny; k++) {
            S.P(i, k, 0) = S.P(i, k, 1) - (S.U(i, k, 1) - 2 * S.U(i, k, 0) + S.U(i, k, -1)) / Re / S.dx;
            S.P(i, k, nx + 1) =
                S.P(i, k, nx) - (S.U(i, k, nx + 1) - 2 * S.U(i, k, nx) + S.U(i, k, nx - 1)) / Re / S.dx;
        }
    for (int i = 1; i <= nz; i++)
        for (int j = 1; j <= nx; j++) {
            S.P(i, 0, j) = S.P(i, 1, j) - (S.V(i, 1, j) - 2 * S.V(i, 0, j) + S.V(i, -1, j)) / Re / S.dy;
            S.P(i, ny + 1, j) =
                S.P(i, ny, j) - (S.V(i, ny + 1, j) - 2 * S.V(i, ny, j) + S.V(i, ny - 1, j)) / Re / S.dy;
        }
    for (int k = 1; k <= ny; k++)
        for (int j = 1; j <= nx; j++) {
            S.P(0, k, j) = S.P(1, k, j) - (S.W(1, k, j) - 2 * S.W(0, k, j) + S.W(-1, k, j)) / Re / S.dz;
            S.P(nz + 1, k, j) =
                S.P(nz, k, j) - (S.W(nz + 1, k, j) - 2 * S.W(nz, k, j) + S.W(nz - 1, k, j)) / Re / S.dz;
        }
}

// FGH (src/ns_cube.cpp:131-201)
void ns_fgh(NsState& S) {
    const int nx = S.nx, ny = S.ny, nz = S.nz;
    const double Re = S.Re, dt = S.dt, dx = S.dx, dy = S.dy, dz = S.dz;
    const double dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 1; i <= nz; i++)
        for (int k = 1; k <= ny; k++)
            for (int j = 0; j <= nx; j++) {
                const double uc = S.U(i, k, j);
                S.Fv(i, k, j) =
                    uc + dt * ((S.U(i, k, j + 1) - 2 * uc + S.U(i, k, j - 1)) / Re / dx2 +
                               (S.U(i, k + 1, j) - 2 * uc + S.U(i, k - 1, j)) / Re / dy2 +
                               (S.U(i + 1, k, j) - 2 * uc + S.U(i - 1, k, j)) / Re / dz2 -
                               (sqd(0.5 * (uc + S.U(i, k, j + 1))) - sqd(0.5 * (S.U(i, k, j - 1) + uc))) / dx -
                               0.25 * ((uc + S.U(i, k + 1, j)) * (S.V(i, k, j + 1) + S.V(i, k, j)) -
                                       (S.U(i, k - 1, j) + uc) * (S.V(i, k - 1, j + 1) + S.V(i, k - 1, j))) / dy -
                               0.25 * ((uc + S.U(i + 1, k, j)) * (S.W(i, k, j + 1) + S.W(i, k, j)) -
                                       (S.U(i - 1, k, j) + uc) * (S.W(i - 1, k, j + 1) + S.W(i - 1, k, j))) / dz);
            }
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 1; i <= nz; i++)
        for (int k = 0; k <= ny; k++)
            for (int j = 1; j <= nx; j++) {
                const double vc = S.V(i, k, j);
                S.Gv(i, k, j) =
                    vc + dt * ((S.V(i, k, j + 1) - 2 * vc + S.V(i, k, j - 1)) / Re / dx2 +
                               (S.V(i, k + 1, j) - 2 * vc + S.V(i, k - 1, j)) / Re / dy2 +
                               (S.V(i + 1, k, j) - 2 * vc + S.V(i - 1, k, j)) / Re / dz2 -
                               (sqd(0.5 * (vc + S.V(i, k + 1, j))) - sqd(0.5 * (S.V(i, k - 1, j) + vc))) / dy -
                               0.25 * ((S.U(i, k, j) + S.U(i, k + 1, j)) * (S.V(i, k, j + 1) + vc) -
                                       (S.U(i, k, j - 1) + S.U(i, k + 1, j - 1)) * (vc + S.V(i, k, j - 1))) / dx -
                               0.25 * ((S.W(i, k, j) + S.W(i, k + 1, j)) * (vc + S.V(i + 1, k, j)) -
                                       (S.W(i - 1, k, j) + S.W(i - 1, k + 1, j)) * (S.V(i - 1, k, j) + vc)) / dz);
            }
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i <= nz; i++)
        for (int k = 1; k <= ny; k++)
            for (int j = 1; j <= nx; j++) {
                const double wc = S.W(i, k, j);
                S.Hv(i, k, j) =
                    wc + dt * ((S.W(i, k, j + 1) - 2 * wc + S.W(i, k, j - 1)) / Re / dx2 +
                               (S.W(i, k + 1, j) - 2 * wc + S.W(i, k - 1, j)) / Re / dy2 +
                               (S.W(i + 1, k, j) - 2 * wc + S.W(i - 1, k, j)) / Re / dz2 -
                               (sqd(0.5 * (S.W(i + 1, k, j) + wc)) - sqd(0.5 * (S.W(i - 1, k, j) + wc))) / dz -
                               0.25 * ((S.U(i + 1, k, j) + S.U(i, k, j)) * (S.W(i, k, j + 1) + wc) -
                                       (S.U(i + 1, k, j - 1) + S.U(i, k, j - 1)) * (wc + S.W(i, k, j - 1))) / dx -
                               0.25 * ((wc + S.W(i, k + 1, j)) * (S.V(i, k, j) + S.V(i + 1, k, j)) -
                                       (S.W(i, k - 1, j) + wc) * (S.V(i, k - 1, j) + S.V(i + 1, k - 1, j))) / dy);
            }
}

template <typename T>
int lapl_cube_solve(T* ans, const T* rhs, int periodic, double dx, double dy, double dz,
                    double lx, double ly, double lz, int nx, int ny, int nz);

// poisson + update_uvwp (src/ns_cube.cpp:204-277); p = x copies the index
// intersection, i.e. p's interior (src/tensor.h:104-111, 282-285).
int ns_poisson_update(NsState& S) {
    const int nx = S.nx, ny = S.ny, nz = S.nz;
    ns_poisson_rhs<double>(S.RHS.data(), S.F.data(), S.G.data(), S.H.data(), S.p.data(), nx, ny, nz, S.dx, S.dy,
                           S.dz, S.dt);
    int rc = lapl_cube_solve<double>(S.x.data(), S.RHS.data(), 0, S.dx, S.dy, S.dz, S.lx, S.ly, S.lz, nx, ny, nz);
    if (rc) return rc;
    const double dt = S.dt;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 1; i <= nz; i++)
        for (int k = 1; k <= ny; k++)
            for (int j = 1; j < nx; j++) S.U(i, k, j) = S.Fv(i, k, j) - dt / S.dx * (S.X(i, k, j + 1) - S.X(i, k, j));
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 1; i <= nz; i++)
        for (int k = 1; k < ny; k++)
            for (int j = 1; j <= nx; j++) S.V(i, k, j) = S.Gv(i, k, j) - dt / S.dy * (S.X(i, k + 1, j) - S.X(i, k, j));
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 1; i < nz; i++)
        for (int k = 1; k <= ny; k++)
            for (int j = 1; j <= nx; j++) S.W(i, k, j) = S.Hv(i, k, j) - dt / S.dz * (S.X(i + 1, k, j) - S.X(i, k, j));
    for (int i = 1; i <= nz; i++)
        for (int k = 1; k <= ny; k++)
            for (int j = 1; j <= nx; j++) S.P(i, k, j) = S.X(i, k, j);
    return 0;
}

// ---------------------------------------------------------------------------
// Synthetic inputs (workload definition, not a reference function):
// counter-based splitmix64, value(seed, idx) = 2*(x >> 11)*2^-53 - 1.
// ---------------------------------------------------------------------------
inline double splitmix_value(uint64_t seed, uint64_t idx) {
    uint64_t z = seed * 0x9E3779B97F4A7C15ull + (idx + 1) * 0xD1B54A32D192ED03ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return 2.0 * (double)(z >> 11) * 0x1.0p-53 - 1.0;
}

}  // namespace

// ---------------------------------------------------------------------------
// LaplCyl3FFT2<T, check, zflag, use_cyclic_reduction>::solve
// (src/lapl_cyl.cpp:11-128) with LaplCyl3Data (src/lapl_cyl.h:12-37) and
// init_solver (src/lapl_cyl.cpp:130-170).  Arrays [nphi][nz][nr] (r fastest)
// over the index ranges {0..nphi-1, z1..zn, 1..nr}.  The r systems: gtsv
// (LU with partial pivoting, the gttrf/gttrs pair of the default
// instantiation) or the general cyclic reduction (use_cyclic_reduction).
// ---------------------------------------------------------------------------
template <typename T>
int lapl_cyl3_fft2_solve(T* ans, const T* rhs, int zper, double dr, double dz, double r0, double lz, int nr,
                         int nz, int nphi, int use_cr) {
    constexpr T SQRT_M_1_PI = (T)0.56418958354775629;
    const double dphi = 2 * M_PI / nphi, dr2 = dr * dr, dz2 = dz * dz, dphi2 = dphi * dphi;
    const double slz = std::sqrt(2. / lz);
    const int zpoints = zper ? nz : nz + 1, z1 = zper ? 0 : 1, zn = zper ? nz - 1 : nz;
    const int nrq = (int)std::ceil(std::log2(nr + 1));
    if (!is_pow2(nphi) || !is_pow2(zpoints) || nr < 1) return -1;
    std::vector<T> lm_phi(nphi), lm_z(zpoints);
    for (int i = 0; i < nphi; i++) {
        const double s = std::sin(i * dphi * 0.5);
        lm_phi[i] = (T)(4.0 / dphi2 * (s * s));
    }
    for (int k = 0; k < zpoints; k++) {
        const double s = zper ? std::sin(k * M_PI / zpoints) : std::sin(k * M_PI * 0.5 / zpoints);
        lm_z[k] = (T)(4. / dz2 * (s * s));
    }
    const int zc = zn - z1 + 1;
    auto at = [&](int i, int k, int j) { return ((size_t)i * zc + (k - z1)) * nr + (j - 1); };
    std::vector<T> W((size_t)nphi * zc * nr);
    const int maxp = std::max(nphi, zpoints);
    auto phi_pass = [&](const T* src, T* dst, T factor, bool inverse) {
        CFFT<T> f(nphi);
#pragma omp parallel
        {
            std::vector<T> sa(maxp + 1), Sa(maxp + 1);
            std::vector<std::complex<T>> buf(2 * maxp + 2);
#pragma omp for collapse(2) schedule(static)
            for (int k = z1; k <= zn; k++)
                for (int j = 1; j <= nr; j++) {
                    for (int i = 0; i < nphi; i++) sa[i] = src[at(i, k, j)];
                    if (!inverse) rdft_fwd<T>(f, nphi, sa.data(), Sa.data(), factor, buf.data());
                    else rdft_inv<T>(f, nphi, sa.data(), Sa.data(), factor, buf.data());
                    for (int i = 0; i < nphi; i++) dst[at(i, k, j)] = Sa[i];
                }
        }
    };
    auto z_pass = [&](const T* src, T* dst, T factor, bool inverse) {
        CFFT<T> f(zper ? zpoints : 2 * zpoints);
#pragma omp parallel
        {
            std::vector<T> sa(maxp + 1), Sa(maxp + 1);
            std::vector<std::complex<T>> buf(2 * maxp + 2);
#pragma omp for collapse(2) schedule(static)
            for (int i = 0; i < nphi; i++)
                for (int j = 1; j <= nr; j++) {
                    for (int k = z1; k <= zn; k++) sa[k] = src[at(i, k, j)];
                    if (!zper) dst_pair<T>(f, zpoints, sa.data(), nullptr, Sa.data(), nullptr, factor, buf.data());
                    else if (!inverse) rdft_fwd<T>(f, zpoints, sa.data(), Sa.data(), factor, buf.data());
                    else rdft_inv<T>(f, zpoints, sa.data(), Sa.data(), factor, buf.data());
                    for (int k = z1; k <= zn; k++) dst[at(i, k, j)] = Sa[k];
                }
        }
    };
    phi_pass(rhs, W.data(), (T)(dphi * SQRT_M_1_PI), false);
    z_pass(W.data(), W.data(), (T)(dz * slz), false);
    int info_all = 0;
#pragma omp parallel
    {
        std::vector<T> L(nr), D(nr), U(nr), b(nr);
#pragma omp for collapse(2) schedule(static)
        for (int i = 0; i < nphi; i++)
            for (int k = z1; k <= zn; k++) {
                for (int j = 1; j <= nr; j++) {
                    const double r = r0 + j * dr;
                    D[j - 1] = (T)(-2 / dr2 - lm_phi[i] / r / r - lm_z[k]);
                    L[j - 1] = (T)((r - 0.5 * dr) / dr2 / r);   // row j's sub-diagonal (unused at j = 1)
                    U[j - 1] = (T)((r + 0.5 * dr) / dr2 / r);   // row j's super-diagonal (unused at j = nr)
                    b[j - 1] = W[at(i, k, j)];
                }
                if (use_cr) {
                    std::vector<T> e(nr), f(nr);
                    for (int j = 0; j < nr; j++) {
                        e[j] = j > 0 ? L[j] : T(0);
                        f[j] = j < nr - 1 ? U[j] : T(0);
                    }
                    cyclic_reduction<T>(D.data(), e.data(), f.data(), b.data(), nrq, nr, true);
                } else {
                    // gtsv: dl[j] = A(j+1, j) (row j+1's sub-diagonal), du[j] = A(j, j+1)
                    std::vector<T> dl(nr), du(nr);
                    for (int j = 0; j + 1 < nr; j++) {
                        dl[j] = L[j + 1];
                        du[j] = U[j];
                    }
                    const int info = gtsv<T>(nr, dl.data(), D.data(), du.data(), b.data());
                    if (info) {
#pragma omp critical
                        info_all = info;
                    }
                }
                for (int j = 1; j <= nr; j++) W[at(i, k, j)] = b[j - 1];
            }
    }
    if (info_all) return info_all;
    z_pass(W.data(), W.data(), (T)slz, true);
    phi_pass(W.data(), ans, SQRT_M_1_PI, true);
    return 0;
}

extern "C" {

// ----- transforms ----------------------------------------------------------
// PINNED-BY: test_dst_vs_direct, test_dst_vs_scipy, test_dst_vs_ref, golden sfft vectors.
void orc_dst1_d(double* X, const double* x, int N, double scale) {
    CFFT<double> f(2 * N);
    std::vector<std::complex<double>> buf(2 * N);
    dst_pair<double>(f, N, x, nullptr, X, nullptr, scale, buf.data());
}
void orc_dst1_f(float* X, const float* x, int N, float scale) {
    CFFT<float> f(2 * N);
    std::vector<std::complex<float>> buf(2 * N);
    dst_pair<float>(f, N, x, nullptr, X, nullptr, scale, buf.data());
}
// PINNED-BY: test_dct_vs_direct, test_dct_vs_ref.
void orc_dct1_d(double* X, const double* x, int N, double scale) {
    CFFT<double> f(2 * N);
    std::vector<std::complex<double>> buf(2 * N);
    dct1<double>(f, N, x, X, scale, buf.data());
}
// PINNED-BY: test_rdft_vs_direct, test_rdft_roundtrip, test_rdft_vs_ref.
void orc_rdft_fwd_d(double* X, const double* x, int N, double scale) {
    CFFT<double> f(N);
    std::vector<std::complex<double>> buf(N);
    rdft_fwd<double>(f, N, x, X, scale, buf.data());
}
void orc_rdft_inv_d(double* x, const double* X, int N, double scale) {
    CFFT<double> f(N);
    std::vector<std::complex<double>> buf(N);
    rdft_inv<double>(f, N, X, x, scale, buf.data());
}
// Slow direct sums in long double (the shape of sFT / cFT,
// src/asp_fft.cpp:308-318, 404-416): independent closed-form checkers.
void orc_sft_direct(double* X, const double* x, int N, double scale) {
    for (int k = 1; k < N; k++) {
        long double s = 0;
        for (int j = 1; j < N; j++) s += (long double)x[j] * sinl((long double)M_PI * ((j * k) % (2 * N)) / N);
        X[k] = (double)(s * scale);
    }
}

// ----- solvers --------------------------------------------------------------
// PINNED-BY: test_cube_vs_ref (bitwise-near, oracle/_ref), test_cube_manufactured,
//            test_cube_dense (brute force 7^3), golden cube vectors.
int orc_lapl_cube_d(double* ans, const double* rhs, int periodic, double dx, double dy, double dz,
                    double lx, double ly, double lz, int nx, int ny, int nz) {
    return lapl_cube_solve<double>(ans, rhs, periodic, dx, dy, dz, lx, ly, lz, nx, ny, nz);
}
int orc_lapl_cube_f(float* ans, const float* rhs, int periodic, double dx, double dy, double dz,
                    double lx, double ly, double lz, int nx, int ny, int nz) {
    return lapl_cube_solve<float>(ans, rhs, periodic, dx, dy, dz, lx, ly, lz, nx, ny, nz);
}
// PINNED-BY: test_rect_manufactured, test_rect_biquadratic (ut_lapl_rect.cpp:300-374),
//            test_rect_vs_fft2 (ut_lapl_rect.cpp:384-456), test_rect_dense.
int orc_lapl_rect_d(double* ans, const double* rhs, int yper, double dx, double dy, double lx, double ly,
                    int nx, int ny, const double* lm_y_scale, const double* L_scale, const double* U_scale) {
    return lapl_rect_solve<double>(ans, rhs, yper, dx, dy, lx, ly, nx, ny, lm_y_scale, L_scale, U_scale);
}
int orc_lapl_rect_f(float* ans, const float* rhs, int yper, double dx, double dy, double lx, double ly,
                    int nx, int ny, const float* lm_y_scale, const float* L_scale, const float* U_scale) {
    return lapl_rect_solve<float>(ans, rhs, yper, dx, dy, lx, ly, nx, ny, lm_y_scale, L_scale, U_scale);
}
// PINNED-BY: test_rect_vs_fft2, test_fft2_dense, test_fft2_periodic_operator.
int orc_lapl_rect_fft2_d(double* ans, const double* rhs, int yper, int xper, double dx, double dy,
                         double lx, double ly, int nx, int ny, const double* lm_y_scale) {
    return lapl_rect_fft2_solve<double>(ans, rhs, yper, xper, dx, dy, lx, ly, nx, ny, lm_y_scale);
}

// PINNED-BY: test_cyl3_dense, test_cyl3_operator, test_cyl3_known_answer_second_order
//            (ut/ut_lapl_cyl.cpp:320-406).
int orc_lapl_cyl3_fft2_d(double* ans, const double* rhs, int zper, double dr, double dz, double r0, double lz,
                         int nr, int nz, int nphi, int use_cr) {
    return lapl_cyl3_fft2_solve<double>(ans, rhs, zper, dr, dz, r0, lz, nr, nz, nphi, use_cr);
}
int orc_lapl_cyl3_fft2_f(float* ans, const float* rhs, int zper, double dr, double dz, double r0, double lz,
                         int nr, int nz, int nphi, int use_cr) {
    return lapl_cyl3_fft2_solve<float>(ans, rhs, zper, dr, dz, r0, lz, nr, nz, nphi, use_cr);
}

// ----- tridiagonal ----------------------------------------------------------
// PINNED-BY: test_gtsv_vs_scipy.
int orc_gtsv_d(int n, double* dl, double* d, double* du, double* b) { return gtsv<double>(n, dl, d, du, b); }
int orc_gtsv_f(int n, float* dl, float* d, float* du, float* b) { return gtsv<float>(n, dl, d, du, b); }
// PINNED-BY: test_thomas_vs_dense, test_thomas_vs_ref.
void orc_thomas_d(double* B, double* A1, double* A2, double* A3, int n) { thomas<double>(B, A1, A2, A3, n); }
// PINNED-BY: test_cr_vs_dense, test_cr_vs_ref.
void orc_cr_d(double* d, double* e, double* f, double* b, int q, int n, int general) {
    cyclic_reduction<double>(d, e, f, b, q, n, general != 0);
}
void orc_cr_f(float* d, float* e, float* f, float* b, int q, int n, int general) {
    cyclic_reduction<float>(d, e, f, b, q, n, general != 0);
}

// ----- caller-side ----------------------------------------------------------
// PINNED-BY: test_ns_rhs_manufactured (divergence of a known field).
void orc_ns_poisson_rhs_d(double* RHS, const double* F, const double* G, const double* H, const double* p,
                          int nx, int ny, int nz, double dx, double dy, double dz, double dt) {
    ns_poisson_rhs<double>(RHS, F, G, H, p, nx, ny, nz, dx, dy, dz, dt);
}

// ----- NSCube (constructor arguments as NSCube(const Config&), src/ns_cube.h:46-78)
void* orc_ns_create(int nx, int ny, int nz, double x1, double x2, double y1, double y2, double z1, double z2,
                    double U0, double Re, double dt) {
    if (nx < 2 || ny < 2 || nz < 2 || nz > nx) return nullptr;
    NsState* S = new NsState();
    S->nx = nx; S->ny = ny; S->nz = nz;
    S->dx = (x2 - x1) / nx; S->dy = (y2 - y1) / ny; S->dz = (z2 - z1) / nz;
    S->lx = x2 - x1 + S->dx; S->ly = y2 - y1 + S->dy; S->lz = z2 - z1 + S->dz;
    S->U0 = U0; S->Re = Re; S->dt = dt;
    S->u.assign((size_t)(nz + 2) * (ny + 2) * (nx + 3), 0.0);
    S->v.assign((size_t)(nz + 2) * (ny + 3) * (nx + 2), 0.0);
    S->w.assign((size_t)(nz + 3) * (ny + 2) * (nx + 2), 0.0);
    S->p.assign((size_t)(nz + 2) * (ny + 2) * (nx + 2), 0.0);
    S->x.assign((size_t)nz * ny * nx, 0.0);
    S->RHS.assign((size_t)nz * ny * nx, 0.0);
    S->F.assign((size_t)nz * ny * (nx + 1), 0.0);
    S->G.assign((size_t)nz * (ny + 1) * nx, 0.0);
    S->H.assign((size_t)(nz + 1) * ny * nx, 0.0);
    return S;
}

// NSCube::step() (src/ns_cube.cpp:27-62) `count` times; returns 0 or the solver's error.
int orc_ns_step(void* h, int count) {
    NsState* S = (NsState*)h;
    for (int c = 0; c < count; c++) {
        ns_init_bound(*S);
        ns_fgh(*S);
        int rc = ns_poisson_update(*S);
        if (rc) return rc;
        S->steps++;
    }
    return 0;
}

// Copies one array out: 0 u, 1 v, 2 w, 3 p, 4 x, 5 F, 6 G, 7 H, 8 RHS; returns its length.
long orc_ns_get(void* h, int which, double* out) {
    NsState* S = (NsState*)h;
    std::vector<double>* a[] = {&S->u, &S->v, &S->w, &S->p, &S->x, &S->F, &S->G, &S->H, &S->RHS};
    if (which < 0 || which > 8) return -1;
    if (out) std::copy(a[which]->begin(), a[which]->end(), out);
    return (long)a[which]->size();
}

// Overwrites one array (tests start from a given state).
long orc_ns_set(void* h, int which, const double* in) {
    NsState* S = (NsState*)h;
    std::vector<double>* a[] = {&S->u, &S->v, &S->w, &S->p, &S->x, &S->F, &S->G, &S->H, &S->RHS};
    if (which < 0 || which > 8) return -1;
    std::copy(in, in + a[which]->size(), a[which]->begin());
    return (long)a[which]->size();
}

void orc_ns_destroy(void* h) { delete (NsState*)h; }

// ----- workload helpers (inputs and residual norms) ------------------------
// u ~ U[-1,1) on the n points of the interior, seeded.
void orc_field(double* u, uint64_t seed, long n) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) u[i] = splitmix_value(seed, (uint64_t)i);
}

// f = A u, the 7-point (hz > 0) or 5-point (hz == 0, 2-D, nz == 1) discrete
// Laplacian with zero Dirichlet boundary: the operator LaplCube/LaplRect invert.
// hx, hy, hz are the spacings of the operator (see the eigenvalue aliasing
// quirk in lapl_cube_solve).
void orc_apply_lapl(double* f, const double* u, int nx, int ny, int nz, double hx, double hy, double hz) {
    const bool three = hz > 0;
    const double ix = 1 / (hx * hx), iy = 1 / (hy * hy), iz = three ? 1 / (hz * hz) : 0;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < nz; i++)
        for (int k = 0; k < ny; k++)
            for (int j = 0; j < nx; j++) {
                size_t c = ((size_t)i * ny + k) * nx + j;
                double uc = u[c];
                double xm = j > 0 ? u[c - 1] : 0, xp = j < nx - 1 ? u[c + 1] : 0;
                double ym = k > 0 ? u[c - nx] : 0, yp = k < ny - 1 ? u[c + nx] : 0;
                double zm = 0, zp = 0;
                if (three) {
                    zm = i > 0 ? u[c - (size_t)nx * ny] : 0;
                    zp = i < nz - 1 ? u[c + (size_t)nx * ny] : 0;
                }
                f[c] = (xm - 2 * uc + xp) * ix + (ym - 2 * uc + yp) * iy + (three ? (zm - 2 * uc + zp) * iz : 0);
            }
}

// ||f - A u||_2 / ||f||_2 in FP64 with long double accumulation.
double orc_rel_residual(const double* f, const double* u, int nx, int ny, int nz, double hx, double hy,
                        double hz) {
    const size_t n = (size_t)nx * ny * nz;
    std::vector<double> au(n);
    orc_apply_lapl(au.data(), u, nx, ny, nz, hx, hy, hz);
    long double rr = 0, ff = 0;
    for (size_t i = 0; i < n; i++) {
        long double r = (long double)f[i] - au[i];
        rr += r * r;
        ff += (long double)f[i] * f[i];
    }
    return ff > 0 ? (double)sqrtl(rr / ff) : (double)sqrtl(rr);
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
