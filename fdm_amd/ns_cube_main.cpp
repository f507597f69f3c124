// fdm_ns_cube on the MI355X: the reference's NSCube driver
// (test/test_ns_cube.cpp:21-101) written against the drop-in classes
// fdm::NSCube (include/fdm/ns_cube.h) and Config (include/fdm/config.h).
// Configuration as the reference's Config: an INI file (default
// ns_cube.ini; --config=FILE picks another) and command-line overrides
// --section:key=value.  Keys: ns:{x1,y1,z1,x2,y2,z2,u0,Re,dt,nx,nz,steps,
// verbose} (ny reads the key "nx", as NSCube's constructor does,
// src/ns_cube.h:58), plot:{interval,png,vtk}, solver:datatype (double; the
// GPU step keeps FP64 fields), solver:gpu_precision (f64 | refine: the
// pressure solve in FP64 or FP32 with FP64 refinement).  ns:verbose=1 prints
// the reference's per-step maxima line and, at exit, its per-phase step
// times.  VTK snapshots step_%07d.vtk in the format of
// velocity_plotter::vtk_out (src/velocity_plot.cpp:113-218, Cartesian
// branch), and step_%07d.png with the six panels of velocity_plotter::plot
// (velocity_png.hpp: filled 10-level maps, titles in tEXt chunks; the stream
// functions solved on the GPU by the LaplRect drop-ins).  --dry-run prints
// the resolved parameters and exits without touching the GPU.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#define FDMGPU_THROW 1
#include "../include/fdm/config.h"
#include "../include/fdm/ns_cube.h"
#include "velocity_png.hpp"

namespace {

// Cell-centred velocity (the average of the two faces of each staggered
// component) on the interior, as velocity_plotter::vtk_out's Cartesian branch.
template <typename NS>
void vtk_out(const std::string& name, NS& ns) {
    const int nx = ns.nx, ny = ns.ny, nz = ns.nz;
    const std::vector<double> u = ns.u.get(), v = ns.v.get(), w = ns.w.get();
    auto U = [&](int i, int k, int j) { return u[ns.u.offset(i, k, j)]; };
    auto V = [&](int i, int k, int j) { return v[ns.v.offset(i, k, j)]; };
    auto W = [&](int i, int k, int j) { return w[ns.w.offset(i, k, j)]; };
    FILE* f = std::fopen(name.c_str(), "wb");
    if (!f) {
        std::perror(name.c_str());
        std::exit(1);
    }
    std::fprintf(f, "# vtk DataFile Version 3.0\nstep %d\nASCII\n", ns.time_index);
    std::fprintf(f, "DATASET STRUCTURED_POINTS\nDIMENSIONS %d %d %d\nASPECT_RATIO 1 1 1\n", nx, ny, nz);
    std::fprintf(f, "ORIGIN %f %f %f\nSPACING %f %f %f\n", ns.x1, ns.y1, ns.z1, ns.dx, ns.dy, ns.dz);
    std::fprintf(f, "POINT_DATA %d\nVECTORS u double\n", nx * ny * nz);
    for (int i = 1; i <= nz; i++)
        for (int k = 1; k <= ny; k++)
            for (int j = 1; j <= nx; j++)
                std::fprintf(f, "%f %f %f\n", 0.5 * (U(i, k, j) + U(i, k, j - 1)), 0.5 * (V(i, k, j) + V(i, k - 1, j)),
                             0.5 * (W(i, k, j) + W(i - 1, k, j)));
    std::fclose(f);
}

int calc(const Config& c) {
    fdm::NSCube<double, false> ns(c);
    const int steps = c.get("ns", "steps", 1);
    const int plot_interval = c.get("plot", "interval", 100);
    const int png = c.get("plot", "png", 1), vtk = c.get("plot", "vtk", 0);
    std::unique_ptr<fdmpng::VelocityPng> plot;
    if (png)
        plot.reset(new fdmpng::VelocityPng(ns.dx, ns.dy, ns.dz, ns.nx, ns.ny, ns.nz, ns.x1, ns.x2, ns.y1, ns.y2,
                                           ns.z1, ns.z2));
    auto snapshot = [&]() {
        char name[64];
        if (png) {
            std::snprintf(name, sizeof name, "step_%07d.png", ns.time_index);
            if (!plot->plot(name, ns.time_index * ns.dt, ns.u.get(), ns.v.get(), ns.w.get())) {
                std::perror(name);
                std::exit(1);
            }
        }
        if (vtk) {
            std::snprintf(name, sizeof name, "step_%07d.vtk", ns.time_index);
            vtk_out(name, ns);
        }
    };
    if (png || vtk) snapshot();
    const auto t1 = std::chrono::steady_clock::now();
    for (int i = 0; i < steps; i++) {
        ns.step();
        if ((i + 1) % plot_interval == 0 && (png || vtk)) snapshot();
    }
    (void)ns.solver_stats();   // waits for the last step
    const auto t2 = std::chrono::steady_clock::now();
    std::printf("It took me '%f' seconds\n", std::chrono::duration<double>(t2 - t1).count());
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    std::string ini = "ns_cube.ini";
    bool dry = false;
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--dry-run")) dry = true;
        if (!std::strncmp(argv[i], "--config=", 9)) ini = argv[i] + 9;
    }
    Config c;
    c.open(ini);
    c.rewrite(argc, argv);
    const std::string datatype = c.get("solver", "datatype", "double");
    const std::string gp = c.get("solver", "gpu_precision", "f64");
    if (dry) {
        const double x1 = c.get("ns", "x1", -M_PI), y1 = c.get("ns", "y1", -M_PI), z1 = c.get("ns", "z1", -M_PI);
        const double x2 = c.get("ns", "x2", M_PI), y2 = c.get("ns", "y2", M_PI), z2 = c.get("ns", "z2", M_PI);
        std::printf("nx %d ny %d nz %d x [%g, %g] y [%g, %g] z [%g, %g] u0 %g Re %g dt %g steps %d verbose %d\n",
                    c.get("ns", "nx", 32), c.get("ns", "nx", 32), c.get("ns", "nz", 32), x1, x2, y1, y2, z1, z2,
                    c.get("ns", "u0", 1.0), c.get("ns", "Re", 1.0), c.get("ns", "dt", 0.001), c.get("ns", "steps", 1),
                    c.get("ns", "verbose", 0));
        std::printf("plot interval %d png %d vtk %d datatype %s gpu_precision %s\n", c.get("plot", "interval", 100),
                    c.get("plot", "png", 1), c.get("plot", "vtk", 0), datatype.c_str(),
                    gp == "refine" ? "refine" : "f64");
        return 0;
    }
    if (datatype != "double") {
        std::fprintf(stderr, "solver:datatype=%s: the GPU NSCube step keeps FP64 fields; use "
                             "solver:gpu_precision=refine for FP32 transforms in the pressure solve\n",
                     datatype.c_str());
        return 2;
    }
    try {
        return calc(c);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
}
