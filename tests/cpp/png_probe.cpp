// Writes a PNG through fdm_amd/velocity_png.hpp's encoder (no GPU): six panels
// of known fields, for tests/test_png.py to decode and check pixel by pixel.
//   png_probe OUT.png
// Panel k (rows x cols = (7 + k) x (11 + 2k)): value (r * cols + c) * (k + 1)
// - 3k, so each panel spans its own range; titles "title k".
#include <cstdio>
#include <string>
#include <vector>

#include "../../fdm_amd/velocity_png.hpp"

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    std::vector<fdmpng::Field2> panels;
    std::vector<std::string> titles;
    for (int k = 0; k < 6; k++) {
        fdmpng::Field2 f;
        f.rows = 7 + k;
        f.cols = 11 + 2 * k;
        f.v.resize((size_t)f.rows * f.cols);
        for (int r = 0; r < f.rows; r++)
            for (int c = 0; c < f.cols; c++) f.at(r, c) = (r * f.cols + c) * (k + 1.0) - 3.0 * k;
        panels.push_back(f);
        titles.push_back("title " + std::to_string(k));
    }
    return fdmpng::write_panels(argv[1], panels, titles) ? 0 : 1;
}
