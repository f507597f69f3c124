// PNG snapshots of the NSCube driver, after the reference's
// velocity_plotter::update / plot (src/velocity_plot.cpp:18-110): six panels
// in a 2 x 3 grid, each a 10-level filled map of one 2-D field
//   U on the z = nz/2 slice, V on z = nz/2, W on y = ny/2,
//   the stream functions psi_z (x-y slice), psi_y (x-z slice), psi_x (y-z slice),
// where psi solves the 2-D Poisson equation with the slice's vorticity
// (central differences of the cell-centred velocities) on the right-hand side
// -- through this build's LaplRect / LaplRectFFT2 drop-ins, so on the GPU.
// The reference draws with plplot (axes, labels, contour lines); this writer
// has no font: each panel's title (as the reference formats it) is stored in
// a tEXt chunk "Panel<k>" of the PNG instead.  Level colours: 10 equal bins
// between the panel's min and max, a blue-to-red ramp (kPalette).
//
// The PNG is RGB8, zlib stream of stored (uncompressed) deflate blocks, so
// the writer needs no library.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../include/fdm/lapl_rect.h"

namespace fdmpng {

constexpr int kLevels = 10;
constexpr unsigned char kPalette[kLevels][3] = {{49, 54, 149},  {69, 117, 180},  {116, 173, 209}, {171, 217, 233},
                                                {224, 243, 248}, {254, 224, 144}, {253, 174, 97},  {244, 109, 67},
                                                {215, 48, 39},   {165, 0, 38}};
constexpr int kPanel = 256, kGap = 4, kCols = 3, kRows = 2;
constexpr int kWidth = kCols * kPanel + (kCols + 1) * kGap, kHeight = kRows * kPanel + (kRows + 1) * kGap;

// a 2-D field: rows (slow index) x cols, row 0 drawn at the bottom
struct Field2 {
    int rows = 0, cols = 0;
    std::vector<double> v;
    double& at(int r, int c) { return v[(size_t)r * cols + c]; }
    double at(int r, int c) const { return v[(size_t)r * cols + c]; }
    double maxabs() const {
        double m = 0.0;
        for (double t : v) m = m < std::fabs(t) ? std::fabs(t) : m;
        return m;
    }
};

// level of value t in [lo, hi] (10 equal bins; a constant field is level 0)
inline int level(double t, double lo, double hi) {
    if (!(hi > lo)) return 0;
    const int l = (int)std::floor((t - lo) / (hi - lo) * kLevels);
    return l < 0 ? 0 : l >= kLevels ? kLevels - 1 : l;
}

// grid cell of panel pixel p (0 .. kPanel-1) over n cells
inline int cell(int p, int n) { return (int)((long)p * n / kPanel); }

inline uint32_t crc32(const unsigned char* d, size_t n, uint32_t c = 0xffffffffu) {
    for (size_t i = 0; i < n; i++) {
        c ^= d[i];
        for (int k = 0; k < 8; k++) c = (c >> 1) ^ (0xedb88320u & (0u - (c & 1u)));
    }
    return c;
}

inline void put32(std::vector<unsigned char>& o, uint32_t v) {
    for (int s = 24; s >= 0; s -= 8) o.push_back((unsigned char)(v >> s));
}

inline void chunk(std::vector<unsigned char>& png, const char* type, const std::vector<unsigned char>& data) {
    put32(png, (uint32_t)data.size());
    std::vector<unsigned char> td(type, type + 4);
    td.insert(td.end(), data.begin(), data.end());
    png.insert(png.end(), td.begin(), td.end());
    put32(png, crc32(td.data(), td.size()) ^ 0xffffffffu);
}

// panels[k] drawn at grid position (k / 3, k % 3), titles[k] as tEXt "Panel<k>"
inline bool write_panels(const std::string& name, const std::vector<Field2>& panels,
                         const std::vector<std::string>& titles) {
    std::vector<unsigned char> img((size_t)kHeight * (1 + 3 * kWidth), 255);
    for (int y = 0; y < kHeight; y++) img[(size_t)y * (1 + 3 * kWidth)] = 0;   // filter byte: none
    for (size_t k = 0; k < panels.size() && k < (size_t)(kRows * kCols); k++) {
        const Field2& f = panels[k];
        const auto mm = std::minmax_element(f.v.begin(), f.v.end());
        const double lo = *mm.first, hi = *mm.second;
        const int x0 = kGap + (int)(k % kCols) * (kPanel + kGap), y0 = kGap + (int)(k / kCols) * (kPanel + kGap);
        for (int py = 0; py < kPanel; py++) {
            const int r = cell(kPanel - 1 - py, f.rows);
            unsigned char* row = &img[(size_t)(y0 + py) * (1 + 3 * kWidth) + 1 + 3 * x0];
            for (int px = 0; px < kPanel; px++) {
                const unsigned char* c = kPalette[level(f.at(r, cell(px, f.cols)), lo, hi)];
                row[3 * px] = c[0];
                row[3 * px + 1] = c[1];
                row[3 * px + 2] = c[2];
            }
        }
    }
    // zlib: header, stored blocks of at most 65535 bytes, Adler-32
    std::vector<unsigned char> z = {0x78, 0x01};
    for (size_t p = 0; p < img.size();) {
        const size_t n = std::min<size_t>(65535, img.size() - p);
        z.push_back(p + n == img.size() ? 1 : 0);
        z.push_back((unsigned char)(n & 0xff));
        z.push_back((unsigned char)(n >> 8));
        z.push_back((unsigned char)(~n & 0xff));
        z.push_back((unsigned char)((~n >> 8) & 0xff));
        z.insert(z.end(), img.begin() + p, img.begin() + p + n);
        p += n;
    }
    uint32_t a = 1, b = 0;
    for (unsigned char ch : img) {
        a = (a + ch) % 65521u;
        b = (b + a) % 65521u;
    }
    put32(z, (b << 16) | a);
    std::vector<unsigned char> png = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::vector<unsigned char> ihdr;
    put32(ihdr, kWidth);
    put32(ihdr, kHeight);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});   // 8-bit RGB, no interlace
    chunk(png, "IHDR", ihdr);
    for (size_t k = 0; k < titles.size(); k++) {
        const std::string kw = "Panel" + std::to_string(k);
        std::vector<unsigned char> t(kw.begin(), kw.end());
        t.push_back(0);
        t.insert(t.end(), titles[k].begin(), titles[k].end());
        chunk(png, "tEXt", t);
    }
    chunk(png, "IDAT", z);
    chunk(png, "IEND", {});
    FILE* f = std::fopen(name.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(png.data(), 1, png.size(), f) == png.size();
    return std::fclose(f) == 0 && ok;
}

// Host copies of the staggered velocity arrays (index ranges of NSCube:
// u [0,nz+1]x[0,ny+1]x[-1,nx+1], v [0,nz+1]x[-1,ny+1]x[0,nx+1],
// w [-1,nz+1]x[0,ny+1]x[0,nx+1]) -> the six panels, as
// velocity_plotter::update (src/velocity_plot.cpp:18-64) and plot (:67-110).
class VelocityPng {
public:
    VelocityPng(double dx, double dy, double dz, int nx, int ny, int nz, double x1, double x2, double y1, double y2,
                double z1, double z2)
        : nx(nx), ny(ny), nz(nz), dx(dx), dy(dy), dz(dz),
          lapl_x(dy, dz, y2 - y1 + dy, z2 - z1 + dz, ny, nz),
          lapl_y(dx, dz, x2 - x1 + dx, z2 - z1 + dz, nx, nz),
          lapl_z(dx, dy, x2 - x1 + dx, y2 - y1 + dy, nx, ny) {}

    bool plot(const std::string& name, double t, const std::vector<double>& u, const std::vector<double>& v,
              const std::vector<double>& w) {
        auto U = [&](int i, int k, int j) { return u[((size_t)i * (ny + 2) + k) * (nx + 3) + (j + 1)]; };
        auto V = [&](int i, int k, int j) { return v[((size_t)i * (ny + 3) + (k + 1)) * (nx + 2) + j]; };
        auto W = [&](int i, int k, int j) { return w[((size_t)(i + 1) * (ny + 2) + k) * (nx + 2) + j]; };
        const int xm = nx / 2, ym = ny / 2, zm = nz / 2;
        // cell-centred slices over [0, n+1] in both directions
        Field2 vx = make(nz + 2, ny + 2), wx = make(nz + 2, ny + 2);
        for (int i = 0; i <= nz + 1; i++)
            for (int k = 0; k <= ny + 1; k++) {
                vx.at(i, k) = 0.5 * (V(i, k - 1, xm) + V(i, k, xm));
                wx.at(i, k) = 0.5 * (W(i - 1, k, xm) + W(i, k, xm));
            }
        Field2 uy = make(nz + 2, nx + 2), wy = make(nz + 2, nx + 2);
        for (int i = 0; i <= nz + 1; i++)
            for (int j = 0; j <= nx + 1; j++) {
                uy.at(i, j) = 0.5 * (U(i, ym, j - 1) + U(i, ym, j));
                wy.at(i, j) = 0.5 * (W(i - 1, ym, j) + W(i, ym, j));
            }
        Field2 uz = make(ny + 2, nx + 2), vz = make(ny + 2, nx + 2);
        for (int k = 0; k <= ny + 1; k++)
            for (int j = 0; j <= nx + 1; j++) {
                uz.at(k, j) = 0.5 * (U(zm, k, j - 1) + U(zm, k, j));
                vz.at(k, j) = 0.5 * (V(zm, k - 1, j) + V(zm, k, j));
            }
        // stream functions on the interiors: vorticity by central differences
        Field2 rx = make(nz, ny), ry = make(nz, nx), rz = make(ny, nx);
        for (int i = 1; i <= nz; i++)
            for (int k = 1; k <= ny; k++)
                rx.at(i - 1, k - 1) = (wx.at(i, k + 1) - wx.at(i, k - 1)) / 2 / dy - (vx.at(i + 1, k) - vx.at(i - 1, k)) / 2 / dz;
        for (int i = 1; i <= nz; i++)
            for (int j = 1; j <= nx; j++)
                ry.at(i - 1, j - 1) = (wy.at(i, j + 1) - wy.at(i, j - 1)) / 2 / dx - (uy.at(i + 1, j) - uy.at(i - 1, j)) / 2 / dz;
        for (int k = 1; k <= ny; k++)
            for (int j = 1; j <= nx; j++)
                rz.at(k - 1, j - 1) = (vz.at(k, j + 1) - vz.at(k, j - 1)) / 2 / dx - (uz.at(k + 1, j) - uz.at(k - 1, j)) / 2 / dy;
        Field2 px = make(nz, ny), py = make(nz, nx), pz = make(ny, nx);
        lapl_x.solve(px.v.data(), rx.v.data());
        lapl_y.solve(py.v.data(), ry.v.data());
        lapl_z.solve(pz.v.data(), rz.v.data());
        char b[6][96];
        std::snprintf(b[0], sizeof b[0], "U (t=%.1e, |max|=%.1e)", t, uz.maxabs());
        std::snprintf(b[1], sizeof b[1], "V (t=%.1e, |max|=%.1e)", t, vz.maxabs());
        std::snprintf(b[2], sizeof b[2], "W (t=%.1e, |max|=%.1e)", t, wy.maxabs());
        std::snprintf(b[3], sizeof b[3], "UV (Z=const) (%.1e)", t);
        std::snprintf(b[4], sizeof b[4], "UW (Y=const) (%.1e)", t);
        std::snprintf(b[5], sizeof b[5], "VW (X=const) (%.1e)", t);
        return write_panels(name, {uz, vz, wy, pz, py, px}, {b[0], b[1], b[2], b[3], b[4], b[5]});
    }

private:
    static Field2 make(int rows, int cols) {
        Field2 f;
        f.rows = rows;
        f.cols = cols;
        f.v.assign((size_t)rows * cols, 0.0);
        return f;
    }
    const int nx, ny, nz;
    const double dx, dy, dz;
    fdm::LaplRectFFT2<double, false> lapl_x;   // psi_x on the y-z slice (rows z, columns y)
    fdm::LaplRect<double, false> lapl_y;       // psi_y on the x-z slice (rows z, columns x)
    fdm::LaplRect<double, false> lapl_z;       // psi_z on the x-y slice (rows y, columns x)
};

}  // namespace fdmpng
