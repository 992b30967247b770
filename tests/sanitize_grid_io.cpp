// Drives the grid container I/O (psp_amd/csrc/grid_io.cpp) through every
// entry point, built with AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_sanitizers.py: a raw file written and loaded in all six storage
// orders, reordered back, trilinear interpolation up and down, VTK plane and
// volume files, and the error paths (missing file, short file, bad order,
// bad sizes).  Any invalid access or undefined operation aborts the run.
//   sanitize_grid_io directory
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "psp_hip.h"

#define CHECK(cond, what)                                         \
    do {                                                          \
        if (!(cond)) {                                            \
            std::fprintf(stderr, "sanitize_grid_io: %s\n", what); \
            return 1;                                             \
        }                                                         \
    } while (0)

int main(int argc, char** argv)
{
    if (argc < 2) return 2;
    const std::string dir = argv[1];
    const int nx = 7, ny = 5, nz = 4, ns = 2;
    for (int type = PSP_HIP_REAL64; type <= PSP_HIP_COMPLEX128; ++type) {
        const int w = type == PSP_HIP_REAL64 ? 1 : 2;  // doubles per scalar
        const size_t n = (size_t)nx * ny * nz * ns * w;
        std::vector<double> nat(n);
        for (size_t i = 0; i < n; ++i) nat[i] = std::sin(0.37 * (double)i) + 0.01 * (double)i;
        for (int order = PSP_HIP_XYZ; order <= PSP_HIP_ZYX; ++order) {
            std::vector<double> ord(n), back(n), loaded(n);
            CHECK(psp_hip_grid_reorder(nat.data(), PSP_HIP_XYZ, ord.data(), order, nx, ny, nz, ns,
                                       type) == PSP_HIP_OK,
                  "reorder");
            const std::string f = dir + "/grid_" + std::to_string(type) + "_" +
                                  std::to_string(order) + ".bin";
            FILE* fp = std::fopen(f.c_str(), "wb");
            CHECK(fp, "open for write");
            CHECK(std::fwrite(ord.data(), sizeof(double), n, fp) == n, "write");
            std::fclose(fp);
            CHECK(psp_hip_grid_load(f.c_str(), nx, ny, nz, ns, type, order, loaded.data()) ==
                      PSP_HIP_OK,
                  "load");
            CHECK(std::memcmp(loaded.data(), ord.data(), n * sizeof(double)) == 0, "load bytes");
            CHECK(psp_hip_grid_reorder(loaded.data(), order, back.data(), PSP_HIP_XYZ, nx, ny, nz,
                                       ns, type) == PSP_HIP_OK,
                  "reorder back");
            CHECK(std::memcmp(back.data(), nat.data(), n * sizeof(double)) == 0, "round trip");
            for (int z = 0; z < nz; ++z)
                for (int y = 0; y < ny; ++y)
                    for (int x = 0; x < nx; ++x)
                        CHECK(psp_hip_grid_index(nx, ny, nz, order, x, y, z) >= 0, "index");
            CHECK(psp_hip_grid_index(nx, ny, nz, order, nx, 0, 0) == -1, "index out of range");
            // up and down
            std::vector<double> up((size_t)11 * 9 * 6 * ns * w), down((size_t)3 * 2 * 2 * ns * w);
            CHECK(psp_hip_grid_interpolate(ord.data(), nx, ny, nz, up.data(), 11, 9, 6, ns, type,
                                           order) == PSP_HIP_OK,
                  "interpolate up");
            CHECK(psp_hip_grid_interpolate(ord.data(), nx, ny, nz, down.data(), 3, 2, 2, ns, type,
                                           order) == PSP_HIP_OK,
                  "interpolate down");
            std::vector<double> one((size_t)ns * w);
            CHECK(psp_hip_grid_interpolate(ord.data(), nx, ny, nz, one.data(), 1, 1, 1, ns, type,
                                           order) != PSP_HIP_OK,
                  "interpolation to a single plane accepted");
            for (int pt = PSP_HIP_PLANE_XY; pt <= PSP_HIP_PLANE_YZ; ++pt) {
                const std::string b =
                    dir + "/plane_" + std::to_string(type) + std::to_string(order) + std::to_string(pt);
                CHECK(psp_hip_grid_write_plane(ord.data(), nx, ny, nz, ns, type, order, pt, 1,
                                               b.c_str()) == PSP_HIP_OK,
                      "write plane");
            }
            const std::string vb = dir + "/vol_" + std::to_string(type) + std::to_string(order);
            CHECK(psp_hip_grid_write_volume(ord.data(), nx, ny, nz, ns, type, order, vb.c_str()) ==
                      PSP_HIP_OK,
                  "write volume");
        }
        // error paths
        std::vector<double> buf(n);
        CHECK(psp_hip_grid_load((dir + "/missing.bin").c_str(), nx, ny, nz, ns, type, 0,
                                buf.data()) != PSP_HIP_OK,
              "missing file accepted");
        const std::string shortf = dir + "/short.bin";
        FILE* fp = std::fopen(shortf.c_str(), "wb");
        CHECK(fp, "open short");
        std::fwrite(buf.data(), sizeof(double), n / 2, fp);
        std::fclose(fp);
        CHECK(psp_hip_grid_load(shortf.c_str(), nx, ny, nz, ns, type, 0, buf.data()) != PSP_HIP_OK,
              "short file accepted");
        CHECK(psp_hip_grid_reorder(nat.data(), 9, buf.data(), 0, nx, ny, nz, ns, type) !=
                  PSP_HIP_OK,
              "bad order accepted");
        CHECK(psp_hip_grid_interpolate(nat.data(), 0, ny, nz, buf.data(), 2, 2, 2, ns, type, 0) !=
                  PSP_HIP_OK,
              "empty grid accepted");
        CHECK(psp_hip_grid_write_plane(nat.data(), nx, ny, nz, ns, type, 0, 0, nz, "x") !=
                  PSP_HIP_OK,
              "plane out of range accepted");
        CHECK(std::strlen(psp_hip_grid_last_error()) > 0, "error message");
    }
    std::printf("sanitize_grid_io: ok\n");
    return 0;
}
