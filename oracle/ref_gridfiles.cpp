// Test infrastructure only: VirtualFluids' own grid-file readers
// (src/gpu/core/DataStructureInitializer/GridReaderFiles/CoordNeighborGeoV.cpp,
// BoundaryQs.cpp, BoundaryValues.cpp, compiled from the reference sources where
// they lie) print what they read, for the grid-file tests to compare with.
//   ref_gridfiles cng <path> <binary 0|1> <coord 0|1>
//   ref_gridfiles qs <path>
//   ref_gridfiles values <path>
// Output: one value per line, doubles with 17 significant digits.
#include "DataStructureInitializer/GridReaderFiles/BoundaryQs.h"
#include "DataStructureInitializer/GridReaderFiles/BoundaryValues.h"
#include "DataStructureInitializer/GridReaderFiles/CoordNeighborGeoV.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

int main(int argc, char** argv)
{
    if (argc < 3) return 2;
    const std::string mode = argv[1], path = argv[2];
    if (mode == "cng" && argc == 5) {
        const bool binary = std::atoi(argv[3]) != 0, coord = std::atoi(argv[4]) != 0;
        CoordNeighborGeoV r(path, binary, coord);
        std::printf("%u\n", r.getLevel());
        for (unsigned int l = 0; l <= r.getLevel(); ++l) {
            const unsigned int n = r.getSize(l) + 1;
            std::printf("%u\n", r.getSize(l));
            if (coord) {
                std::vector<real> v(n);
                r.initalCoords(v.data(), l);
                for (real x : v) std::printf("%.17g\n", (double)x);
            } else {
                std::vector<unsigned int> v(n);
                r.initalNeighbors(v.data(), l);
                for (unsigned int x : v) std::printf("%u\n", x);
            }
        }
        return 0;
    }
    if (mode == "qs") {
        BoundaryQs r(path, false);
        std::printf("%u\n", r.getLevel());
        for (unsigned int l = 0; l <= r.getLevel(); ++l) {
            const unsigned int n = r.getSize(l);
            std::printf("%u\n", n);
            std::vector<int> idx(n);
            std::vector<real> q((size_t)27 * n);
            std::vector<real*> cols(27);
            for (int d = 0; d < 27; ++d) cols[d] = q.data() + (size_t)d * n;
            if (n) {
                r.setIndex(idx.data(), l);
                r.setValues(cols.data(), l);
            }
            for (unsigned int e = 0; e < n; ++e) {
                std::printf("%d\n", idx[e]);
                for (int d = 0; d < 27; ++d) std::printf("%.17g\n", (double)cols[d][e]);
            }
        }
        return 0;
    }
    if (mode == "values") {
        BoundaryValues r(path);
        std::printf("%s\n%u\n", r.getBoundaryCondition().c_str(), r.getLevel());
        std::vector<std::vector<std::vector<real>>> vals(r.getLevel() + 1);
        // setBoundarys appends values[level][column][entry] behind vals[level][column]
        for (auto& v : vals) v.resize(3);
        r.setBoundarys(vals);
        for (unsigned int l = 0; l <= r.getLevel(); ++l) {
            std::printf("%zu\n", vals[l].size());
            for (const auto& col : vals[l]) {
                std::printf("%zu\n", col.size());
                for (real x : col) std::printf("%.17g\n", (double)x);
            }
        }
        return 0;
    }
    return 2;
}
