// fdm_ns_cube on the MI355X: the reference's NSCube driver
// (test/test_ns_cube.cpp:21-101) over the C ABI.  Configuration as the
// reference's Config (src/config.cpp:78-150): an INI file (default
// ns_cube.ini, sections [name], "key = value", ';' / '#' comments) and
// command-line overrides --section:key=value.  Keys: ns:{x1,y1,z1,x2,y2,z2,
// u0,Re,dt,nx,nz,steps} (ny reads the key "nx", as NSCube's constructor does,
// src/ns_cube.h:58), plot:{interval,png,vtk}, solver:datatype (double; the
// GPU step keeps FP64 fields), solver:gpu_precision (f64 | refine: the
// pressure solve in FP64 or FP32 with FP64 refinement).  VTK snapshots
// step_%07d.vtk in the format of velocity_plotter::vtk_out
// (src/velocity_plot.cpp:113-218, Cartesian branch); PNG plots need plplot
// and are not produced.  --dry-run prints the resolved parameters and exits
// without touching the GPU.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../include/fdmgpu.h"

namespace {

using Section = std::map<std::string, std::string>;
using ConfigData = std::map<std::string, Section>;

void load_ini(const std::string& fn, ConfigData& data) {
    FILE* f = std::fopen(fn.c_str(), "rb");
    if (!f) return;   // a missing file leaves the defaults (as the reference)
    char buf[4096];
    Section* cur = nullptr;
    while (std::fgets(buf, sizeof buf, f)) {
        std::string line(buf);
        while (!line.empty() && (line.back() == '\n' || line.back() == '\r')) line.pop_back();
        if (line.empty() || line[0] == '#' || line[0] == ';') continue;
        if (line[0] == '[') {
            const size_t e = line.find(']');
            cur = &data[line.substr(1, e == std::string::npos ? std::string::npos : e - 1)];
            continue;
        }
        if (!cur) continue;
        const size_t c = line.find(';');
        if (c != std::string::npos) line.resize(c);
        const char* sep = " =\t";
        std::vector<char> tmp(line.begin(), line.end());
        tmp.push_back(0);
        char* save = nullptr;
        char* k = strtok_r(tmp.data(), sep, &save);
        char* v = k ? strtok_r(nullptr, sep, &save) : nullptr;
        if (k && v) (*cur)[k] = v;
    }
    std::fclose(f);
}

void apply_args(int argc, char** argv, ConfigData& data) {
    for (int i = 1; i < argc; ++i) {
        std::string s = argv[i];
        if (s.rfind("--", 0) != 0) continue;
        s = s.substr(2);
        const size_t c = s.find(':');
        if (c == std::string::npos) continue;
        const size_t e = s.find('=', c + 1);
        if (e == std::string::npos) continue;
        data[s.substr(0, c)][s.substr(c + 1, e - c - 1)] = s.substr(e + 1);
    }
}

struct Config {
    ConfigData data;
    std::string get(const char* sec, const char* key, const std::string& def) const {
        auto s = data.find(sec);
        if (s == data.end()) return def;
        auto k = s->second.find(key);
        return k == s->second.end() ? def : k->second;
    }
    double getd(const char* sec, const char* key, double def) const {
        const std::string v = get(sec, key, "");
        return v.empty() ? def : std::atof(v.c_str());
    }
    int geti(const char* sec, const char* key, int def) const {
        const std::string v = get(sec, key, "");
        return v.empty() ? def : std::atoi(v.c_str());
    }
};

int check(int rc, const char* what) {
    if (rc) {
        std::fprintf(stderr, "%s: %s (code %d)\n", what, fdmgpu_strerror(rc), rc);
        std::exit(1);
    }
    return rc;
}

// Cell-centred velocity (the average of the two faces of each staggered
// component) on the interior, as velocity_plotter::vtk_out's Cartesian branch.
void vtk_out(const std::string& name, int step, fdmgpu_ns ns, int nx, int ny, int nz, double x1, double y1,
             double z1, double dx, double dy, double dz) {
    std::vector<double> u((size_t)(nz + 2) * (ny + 2) * (nx + 3)), v((size_t)(nz + 2) * (ny + 3) * (nx + 2)),
        w((size_t)(nz + 3) * (ny + 2) * (nx + 2));
    check(fdmgpu_ns_get(ns, FDMGPU_NS_U, u.data(), u.size()), "ns_get u");
    check(fdmgpu_ns_get(ns, FDMGPU_NS_V, v.data(), v.size()), "ns_get v");
    check(fdmgpu_ns_get(ns, FDMGPU_NS_W, w.data(), w.size()), "ns_get w");
    // index ranges: u [0,nz+1][0,ny+1][-1,nx+1], v [0,nz+1][-1,ny+1][0,nx+1], w [-1,nz+1][0,ny+1][0,nx+1]
    auto U = [&](int i, int k, int j) { return u[((size_t)i * (ny + 2) + k) * (nx + 3) + (j + 1)]; };
    auto V = [&](int i, int k, int j) { return v[((size_t)i * (ny + 3) + (k + 1)) * (nx + 2) + j]; };
    auto W = [&](int i, int k, int j) { return w[((size_t)(i + 1) * (ny + 2) + k) * (nx + 2) + j]; };
    FILE* f = std::fopen(name.c_str(), "wb");
    if (!f) {
        std::perror(name.c_str());
        std::exit(1);
    }
    std::fprintf(f, "# vtk DataFile Version 3.0\nstep %d\nASCII\n", step);
    std::fprintf(f, "DATASET STRUCTURED_POINTS\nDIMENSIONS %d %d %d\nASPECT_RATIO 1 1 1\n", nx, ny, nz);
    std::fprintf(f, "ORIGIN %f %f %f\nSPACING %f %f %f\n", x1, y1, z1, dx, dy, dz);
    std::fprintf(f, "POINT_DATA %d\nVECTORS u double\n", nx * ny * nz);
    for (int i = 1; i <= nz; i++)
        for (int k = 1; k <= ny; k++)
            for (int j = 1; j <= nx; j++)
                std::fprintf(f, "%f %f %f\n", 0.5 * (U(i, k, j) + U(i, k, j - 1)), 0.5 * (V(i, k, j) + V(i, k - 1, j)),
                             0.5 * (W(i, k, j) + W(i - 1, k, j)));
    std::fclose(f);
}

}  // namespace

int main(int argc, char** argv) {
    Config c;
    std::string ini = "ns_cube.ini";
    bool dry = false;
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--dry-run")) dry = true;
        if (!std::strncmp(argv[i], "--config=", 9)) ini = argv[i] + 9;
    }
    load_ini(ini, c.data);
    apply_args(argc, argv, c.data);

    const double x1 = c.getd("ns", "x1", -M_PI), y1 = c.getd("ns", "y1", -M_PI), z1 = c.getd("ns", "z1", -M_PI);
    const double x2 = c.getd("ns", "x2", M_PI), y2 = c.getd("ns", "y2", M_PI), z2 = c.getd("ns", "z2", M_PI);
    const double u0 = c.getd("ns", "u0", 1.0), Re = c.getd("ns", "Re", 1.0), dt = c.getd("ns", "dt", 0.001);
    const int nx = c.geti("ns", "nx", 32), ny = c.geti("ns", "nx", 32), nz = c.geti("ns", "nz", 32);
    const int steps = c.geti("ns", "steps", 1);
    const int plot_interval = c.geti("plot", "interval", 100);
    const int png = c.geti("plot", "png", 1), vtk = c.geti("plot", "vtk", 0);
    const std::string datatype = c.get("solver", "datatype", "double");
    const std::string gp = c.get("solver", "gpu_precision", "f64");
    const int prec = gp == "refine" ? FDMGPU_F32_REFINE64 : FDMGPU_F64;
    const double dx = (x2 - x1) / nx, dy = (y2 - y1) / ny, dz = (z2 - z1) / nz;

    if (dry) {
        std::printf("nx %d ny %d nz %d x [%g, %g] y [%g, %g] z [%g, %g] u0 %g Re %g dt %g steps %d\n", nx, ny, nz, x1,
                    x2, y1, y2, z1, z2, u0, Re, dt, steps);
        std::printf("plot interval %d png %d vtk %d datatype %s gpu_precision %s\n", plot_interval, png, vtk,
                    datatype.c_str(), prec == FDMGPU_F64 ? "f64" : "refine");
        return 0;
    }
    if (datatype != "double") {
        std::fprintf(stderr, "solver:datatype=%s: the GPU NSCube step keeps FP64 fields; use "
                             "solver:gpu_precision=refine for FP32 transforms in the pressure solve\n",
                     datatype.c_str());
        return 2;
    }
    if (png) std::fprintf(stderr, "plot:png needs plplot: no PNG plots are written\n");

    fdmgpu_ns ns;
    check(fdmgpu_ns_create(&ns, prec, nx, ny, nz, x1, x2, y1, y2, z1, z2, u0, Re, dt), "ns_create");
    int time_index = 0;
    auto snapshot = [&]() {
        char name[64];
        std::snprintf(name, sizeof name, "step_%07d.vtk", time_index);
        vtk_out(name, time_index, ns, nx, ny, nz, x1, y1, z1, dx, dy, dz);
    };
    if (vtk) snapshot();
    const auto t1 = std::chrono::steady_clock::now();
    for (int i = 0; i < steps; i++) {
        check(fdmgpu_ns_step(ns, 1, nullptr), "ns_step");
        time_index++;
        if ((i + 1) % plot_interval == 0 && vtk) snapshot();
    }
    fdmgpu_plan solver = nullptr;
    check(fdmgpu_ns_solver(ns, &solver), "ns_solver");
    fdmgpu_stats st;
    check(fdmgpu_get_stats(solver, &st), "get_stats");   // synchronises the last step
    const auto t2 = std::chrono::steady_clock::now();
    std::printf("It took me '%f' seconds\n", std::chrono::duration<double>(t2 - t1).count());
    fdmgpu_ns_destroy(ns);
    return 0;
}
