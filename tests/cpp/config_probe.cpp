// Query harness for the Config class and unixbench_score: built once against
// include/fdm/config.h + include/fdm/unixbench_score.h (tests/cpp/Makefile)
// and, where the reference checkout exists, once against the reference's
// own src/config.cpp + src/unixbench_score.cpp (oracle/Makefile, output in
// oracle/_ref/), so tests/test_config.py can feed both the same INI text,
// arguments and queries and require identical answers.
//
//   config_probe INI [ARGS...] < queries
// One query per line; one answer line each:
//   d SEC KEY DEF   get(SEC, KEY, double DEF)     -> %.17g
//   i SEC KEY DEF   get(SEC, KEY, int DEF)        -> %d
//   s SEC KEY DEF   get(SEC, KEY, std::string DEF)
//   D SEC KEY / I SEC KEY / S SEC KEY             getd / geti / gets
//   p               print(FILE*) of the whole configuration, then "."
//   u V1 V2 ...     fdm::unixbench_score of the values -> %.17g
// A query that throws answers "throw".
#include <cstdio>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "config.h"
#include "unixbench_score.h"

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: config_probe INI [--sec:key=value ...] < queries\n");
        return 2;
    }
    Config c;
    c.open(argv[1]);
    c.rewrite(argc, argv);
    std::string line;
    while (std::getline(std::cin, line)) {
        std::istringstream in(line);
        std::string op, sec, key, def;
        in >> op;
        if (op.empty()) continue;
        try {
            if (op == "p") {
                c.print(stdout);
                std::printf(".\n");
            } else if (op == "u") {
                std::vector<double> v;
                double t;
                while (in >> t) v.push_back(t);
                std::printf("%.17g\n", fdm::unixbench_score(v));
            } else {
                in >> sec >> key >> def;
                if (op == "d") std::printf("%.17g\n", c.get(sec, key, std::stod(def)));
                else if (op == "i") std::printf("%d\n", c.get(sec, key, std::stoi(def)));
                else if (op == "s") std::printf("%s\n", c.get(sec, key, def).c_str());
                else if (op == "D") std::printf("%.17g\n", c.getd(sec, key));
                else if (op == "I") std::printf("%d\n", c.geti(sec, key));
                else if (op == "S") std::printf("%s\n", c.gets(sec, key).c_str());
                else std::printf("bad query\n");
            }
        } catch (const std::exception&) {
            std::printf("throw\n");
        }
        std::fflush(stdout);
    }
    return 0;
}
