// Drop-in replacement for the reference's NSCube (src/ns_cube.h:12-78,
// src/ns_cube.cpp:13-62): the lid-driven cavity built from a Config, stepped
// by step(), its pressure projection the LaplCube solve -- all on the
// MI355X through the fdmgpu_ns_* C ABI (include/fdmgpu.h).
//
//   Config c;  c.open("ns_cube.ini");  c.rewrite(argc, argv);
//   fdm::NSCube<double, false> ns(c);
//   for (int i = 0; i < steps; i++) ns.step();
//
// Configuration keys as the reference's constructor reads them
// (src/ns_cube.h:46-62): ns:x1,y1,z1 (default -pi), ns:x2,y2,z2 (pi), ns:u0
// (1), ns:Re (1), ns:dt (0.001), ns:nx (32; ny reads the key "nx" as well),
// ns:nz (32), ns:verbose (0).  One key of this build: solver:gpu_precision
// "f64" | "refine" -- the pressure solve in FP64 or in FP32 with FP64
// iterative refinement; without it the choice of the LaplCube drop-in
// applies (FDMGPU_DOUBLE_MODE).
//
// The fields live in HBM with the reference's index ranges; u, v, w, p, x,
// F, G, H, RHS are handles (ns_field) that copy to and from host vectors,
// report maxabs() and give the device address.  T = double only: the GPU
// step keeps FP64 fields, and NSCube<float> reports FDMGPU_ERR_UNSUPPORTED
// at construction (through the reference's verify() policy).
//
// verbose != 0: every step runs eagerly with GPU events at its four phases
// (fdmgpu_ns_step_timed) and prints, as the reference, the line
//   dt*time_index |p| |u| |v| |w| |x| |RHS| |F| |G| |H|     ("%.1e " each)
// of maxima of absolute values over whole arrays (fdmgpu_ns_maxabs, reduced
// on the GPU); the destructor prints the phases' unixbench scores
// "Step times (ms): init_bound=.., FGH=.., poisson=.., update_uvwp=..".
// The phase times are GPU time between the events; the reference's are host
// wall-clock time of its CPU loops.  verbose == 0: steps are enqueued
// asynchronously and replayed from a hipGraph (fdmgpu_ns_step).
#pragma once

#include <array>
#include <cmath>
#include <cstdio>
#include <string>
#include <type_traits>
#include <vector>

#include "config.h"
#include "fdmgpu_shim.h"
#include "unixbench_score.h"

namespace fdm {

// One field of an fdmgpu_ns: index bounds {i1, i2, k1, k2, j1, j2} (z, y, x;
// x fastest) and element count as the reference's tensor of the same name.
class ns_field {
public:
    ns_field(fdmgpu_ns* ns, int which, std::array<int, 6> idx) : ns_(ns), which_(which), indices(idx) {}

    const std::array<int, 6>& bounds() const { return indices; }
    size_t size() const {
        return (size_t)(indices[1] - indices[0] + 1) * (indices[3] - indices[2] + 1) * (indices[5] - indices[4] + 1);
    }
    // position of (i, k, j) in the arrays of get()/set()
    size_t offset(int i, int k, int j) const {
        return ((size_t)(i - indices[0]) * (indices[3] - indices[2] + 1) + (k - indices[2])) *
                   (indices[5] - indices[4] + 1) +
               (j - indices[4]);
    }
    // copy to the host (waits for the steps enqueued so far)
    std::vector<double> get() const {
        std::vector<double> h(size());
        FDMGPU_VERIFY(fdmgpu_ns_get(*ns_, which_, h.data(), h.size()));
        return h;
    }
    void set(const std::vector<double>& h) { FDMGPU_VERIFY(fdmgpu_ns_set(*ns_, which_, h.data(), h.size())); }
    double maxabs() const {
        double m[9];
        FDMGPU_VERIFY(fdmgpu_ns_maxabs(*ns_, m));
        return m[which_];
    }
    double* device() const {
        void* p = nullptr;
        FDMGPU_VERIFY(fdmgpu_ns_info(*ns_, which_, nullptr, &p));
        return static_cast<double*>(p);
    }

private:
    fdmgpu_ns* ns_;
    int which_;
    std::array<int, 6> indices;
};

template <typename T, bool check>
class NSCube {
public:
    const double x1, y1, z1;
    const double x2, y2, z2;
    const double U0;   // lid velocity

    const double Re;
    const double dt;

    const int nx, ny, nz;
    int verbose;

    const double dx, dy, dz;
    const double dx2, dy2, dz2;

    ns_field u, v, w;
    ns_field p, x;
    ns_field F, G, H, RHS;

    int time_index = 0;

    // per-step phase times in ms (verbose runs)
    std::vector<double> init_bound_times;
    std::vector<double> fgh_times;
    std::vector<double> poisson_times;
    std::vector<double> update_times;

    explicit NSCube(const Config& c)
        : x1(c.get("ns", "x1", -M_PI)), y1(c.get("ns", "y1", -M_PI)), z1(c.get("ns", "z1", -M_PI)),
          x2(c.get("ns", "x2", M_PI)), y2(c.get("ns", "y2", M_PI)), z2(c.get("ns", "z2", M_PI)),
          U0(c.get("ns", "u0", 1.0)), Re(c.get("ns", "Re", 1.0)), dt(c.get("ns", "dt", 0.001)),
          nx(c.get("ns", "nx", 32)), ny(c.get("ns", "nx", 32)), nz(c.get("ns", "nz", 32)),
          verbose(c.get("ns", "verbose", 0)), dx((x2 - x1) / nx), dy((y2 - y1) / ny), dz((z2 - z1) / nz),
          dx2(dx * dx), dy2(dy * dy), dz2(dz * dz),
          u(&ns_, FDMGPU_NS_U, {0, nz + 1, 0, ny + 1, -1, nx + 1}),
          v(&ns_, FDMGPU_NS_V, {0, nz + 1, -1, ny + 1, 0, nx + 1}),
          w(&ns_, FDMGPU_NS_W, {-1, nz + 1, 0, ny + 1, 0, nx + 1}),
          p(&ns_, FDMGPU_NS_P, {0, nz + 1, 0, ny + 1, 0, nx + 1}), x(&ns_, FDMGPU_NS_X, {1, nz, 1, ny, 1, nx}),
          F(&ns_, FDMGPU_NS_F, {1, nz, 1, ny, 0, nx}), G(&ns_, FDMGPU_NS_G, {1, nz, 0, ny, 1, nx}),
          H(&ns_, FDMGPU_NS_H, {0, nz, 1, ny, 1, nx}), RHS(&ns_, FDMGPU_NS_RHS, {1, nz, 1, ny, 1, nx}) {
        if (!std::is_same<T, double>::value) FDMGPU_VERIFY(FDMGPU_ERR_UNSUPPORTED);
        const std::string gp = c.get("solver", "gpu_precision", "");
        const int prec = gp == "refine" ? FDMGPU_F32_REFINE64 : gp == "f64" ? FDMGPU_F64 : gpu::double_mode_ref();
        FDMGPU_VERIFY(fdmgpu_ns_create(&ns_, prec, nx, ny, nz, x1, x2, y1, y2, z1, z2, U0, Re, dt));
    }

    NSCube(const NSCube&) = delete;
    NSCube& operator=(const NSCube&) = delete;

    ~NSCube() {
        if (verbose) {
            std::printf("Step times (ms): init_bound=%.2f, FGH=%.2f, poisson=%.2f, update_uvwp=%.2f\n",
                        unixbench_score(init_bound_times), unixbench_score(fgh_times),
                        unixbench_score(poisson_times), unixbench_score(update_times));
        }
        if (ns_) fdmgpu_ns_destroy(ns_);
    }

    void step() {
        if (!verbose) {
            FDMGPU_VERIFY(fdmgpu_ns_step(ns_, 1, nullptr));
            time_index++;
            return;
        }
        double ms[4];
        FDMGPU_VERIFY(fdmgpu_ns_step_timed(ns_, ms, nullptr));
        time_index++;
        double m[9];
        FDMGPU_VERIFY(fdmgpu_ns_maxabs(ns_, m));
        std::printf("%.1e %.1e %.1e %.1e %.1e %.1e %.1e %.1e %.1e %.1e \n", dt * time_index, m[FDMGPU_NS_P],
                    m[FDMGPU_NS_U], m[FDMGPU_NS_V], m[FDMGPU_NS_W], m[FDMGPU_NS_X], m[FDMGPU_NS_RHS], m[FDMGPU_NS_F],
                    m[FDMGPU_NS_G], m[FDMGPU_NS_H]);
        init_bound_times.emplace_back(ms[0]);
        fgh_times.emplace_back(ms[1]);
        poisson_times.emplace_back(ms[2]);
        update_times.emplace_back(ms[3]);
    }

    // The pressure solver (the reference's member lapl_solver): statistics,
    // refinement settings, timing through the fdmgpu_plan calls.
    fdmgpu_plan lapl_solver_plan() const {
        fdmgpu_plan pl = nullptr;
        FDMGPU_VERIFY(fdmgpu_ns_solver(ns_, &pl));
        return pl;
    }
    fdmgpu_stats solver_stats() const {
        fdmgpu_stats st;
        FDMGPU_VERIFY(fdmgpu_get_stats(lapl_solver_plan(), &st));
        return st;
    }

    fdmgpu_ns handle() const { return ns_; }

private:
    fdmgpu_ns ns_ = nullptr;
};

}  // namespace fdm
