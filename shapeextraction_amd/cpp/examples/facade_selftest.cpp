// Host-only checks of the drop-in API's CPU parts: Morton host functions
// (known answers, round trip), DataIO readers/writers, PLY output and error
// reporting.  Prints "facade selftest ok" and exits 0 on success.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>

#include "DataIO.hpp"
#include "MortonUtils.hpp"
#include "ShapesError.hpp"

#define CHECK(c)                                                      \
    do {                                                              \
        if (!(c)) {                                                   \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            return 1;                                                 \
        }                                                             \
    } while (0)

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    // Morton: unit cube, x -> bit 0, y -> bit 1, z -> bit 2
    const uint32_t s = (1u << 21) - 1;
    CHECK(pointToCode(0, 0, 0, 0, 1) == 0);
    CHECK(pointToCode(1, 1, 1, 0, 1) == (uint64_t(1) << 63) - 1);
    CHECK(pointToCode(1.0f / s, 0, 0, 0, 1) == 1);
    CHECK(pointToCode(0, 1.0f / s, 0, 0, 1) == 2);
    CHECK(pointToCode(0, 0, 1.0f / s, 0, 1) == 4);
    for (uint64_t c : {uint64_t(0), uint64_t(12345678901234ull), (uint64_t(1) << 63) - 1, uint64_t(0x5555555555ull)}) {
        // on the unit cube decode -> encode is the identity
        Point p = codeToPoint(c, 0.0f, 1.0f);
        CHECK(pointToCode(p.x, p.y, p.z, 0.0f, 1.0f) == c);
    }
    // DataIO round trips
    PointCloud<float> c;
    c.x_vals = {1.5f, -2.0f, 3.25f};
    c.y_vals = {0.0f, 7.0f, -1.0f};
    c.z_vals = {9.0f, 8.0f, 0.5f};
    DataIO::saveKitti(c, dir + "/t.bin");
    PointCloud<float> k = DataIO::loadFile(dir + "/t.bin");
    CHECK(k.x_vals == c.x_vals && k.y_vals == c.y_vals && k.z_vals == c.z_vals);
    {
        std::ofstream f(dir + "/t.obj");
        f << "# comment\nv 1 2 3\nvn 0 0 1\nv -1.5 0 2\nf 1 2 3\n";
    }
    PointCloud<float> o = DataIO::loadFile(dir + "/t.obj");
    CHECK(o.x_vals.size() == 2 && o.x_vals[1] == -1.5f && o.z_vals[0] == 3.0f);
    {
        std::ofstream f(dir + "/t.txt");
        f << "1 2 3 10 20 30\n4 5 6 1 2 3\nbad line\n";
    }
    PointCloud<float> t = DataIO::loadFile(dir + "/t.txt");
    CHECK(t.x_vals.size() == 2 && t.y_vals[1] == 5.0f);
    bool threw = false;
    try {
        DataIO::loadFile(dir + "/t.xyz");
    } catch (const shapes::Error& e) {
        threw = e.status() == SX_EINVAL;
    }
    CHECK(threw);
    c.setNormals({Point(0, 0, 1), Point(0, 1, 0), Point(1, 0, 0)});
    CHECK(c.saveAsPly(dir + "/t.ply") && c.saveAsBinaryPly(dir + "/t.bin.ply"));
    std::ifstream ply(dir + "/t.ply");
    std::string first;
    std::getline(ply, first);
    CHECK(first == "ply");
    std::printf("facade selftest ok\n");
    return 0;
}
