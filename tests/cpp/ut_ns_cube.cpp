// C++ drop-in check of fdm::NSCube (include/fdm/ns_cube.h) used as the
// reference's test/test_ns_cube.cpp uses NSCube: a Config from an INI file
// and --section:key=value arguments, then step() in a loop.
//
//   ut_ns_cube INI [--ns:key=value ...]
// Runs ns:steps steps of a cube built with the configuration as given (the
// test sets ns:verbose=1, so stdout carries the reference's per-step line and
// the destructor's "Step times" line), then checks, against a second cube
// built from the same configuration with verbose=0 (graph-replayed steps):
//   * time_index == steps on both;
//   * every field bitwise equal between the timed eager steps and the
//     replayed ones;
//   * maxabs() of every field equal to the maximum of |value| over its host
//     copy, NaNs skipped, and offset() placing (i, k, j) where the field's
//     index bounds say.
// Ends with "OK (0 failures)" on success.
#define FDMGPU_THROW 1
#include <cmath>
#include <cstdio>
#include <vector>

#include "fdm/config.h"
#include "fdm/ns_cube.h"

static int failures = 0;
#define CHECK(cond, ...)                                    \
    do {                                                    \
        if (!(cond)) {                                      \
            std::printf("FAIL %s:%d ", __FILE__, __LINE__); \
            std::printf(__VA_ARGS__);                       \
            std::printf("\n");                              \
            failures++;                                     \
        }                                                   \
    } while (0)

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: ut_ns_cube INI [--section:key=value ...]\n");
        return 2;
    }
    Config c;
    c.open(argv[1]);
    c.rewrite(argc, argv);
    const int steps = c.get("ns", "steps", 1);
    Config quiet = c;
    char off[] = "--ns:verbose=0";
    char* qargv[] = {argv[0], off};
    quiet.rewrite(2, qargv);
    try {
        using NS = fdm::NSCube<double, false>;
        NS* a = new NS(c);
        NS b(quiet);
        CHECK(b.verbose == 0, "quiet cube verbose %d", b.verbose);
        for (int i = 0; i < steps; i++) a->step();
        for (int i = 0; i < steps; i++) b.step();
        CHECK(a->time_index == steps && b.time_index == steps, "time_index %d %d", a->time_index, b.time_index);
        fdm::ns_field NS::*fields[] = {&NS::u, &NS::v, &NS::w, &NS::p, &NS::x, &NS::F, &NS::G, &NS::H, &NS::RHS};
        const char* names[] = {"u", "v", "w", "p", "x", "F", "G", "H", "RHS"};
        for (int f = 0; f < 9; f++) {
            const fdm::ns_field& fa = (*a).*fields[f];
            const fdm::ns_field& fb = b.*fields[f];
            const std::vector<double> ha = fa.get(), hb = fb.get();
            CHECK(ha.size() == fa.size(), "%s size", names[f]);
            CHECK(ha == hb, "%s: timed eager steps differ from replayed steps", names[f]);
            double m = 0.0;
            for (double t : ha) m = m < std::fabs(t) ? std::fabs(t) : m;
            CHECK(fa.maxabs() == m, "%s maxabs %.17g vs host %.17g", names[f], fa.maxabs(), m);
            const auto& bd = fa.bounds();
            CHECK(fa.offset(bd[0], bd[2], bd[4]) == 0 && fa.offset(bd[1], bd[3], bd[5]) == ha.size() - 1,
                  "%s offset", names[f]);
        }
        const fdmgpu_stats st = a->solver_stats();
        CHECK(st.iterations >= 1, "solver iterations %d", st.iterations);
        delete a;   // prints the Step times line (verbose)
    } catch (const std::exception& e) {
        std::printf("FAIL exception: %s\n", e.what());
        failures++;
    }
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "OK", failures);
    return failures ? 1 : 0;
}
