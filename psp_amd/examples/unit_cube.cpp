// Unit-cube driver on the C++ API, structured like the reference's
// tests/UnitCube.cpp: pick a velocity model, Initialize, solve a block of 4
// right-hand sides (point source, three point sources, Gaussian beam, plane
// wave; UnitCube.cpp:485-542) with GMRES(20) to 1e-5, report.
//
//   unit_cube [n=40] [omega=30] [model=2] [fp32=0] [prefix]
// models: 0 uniform, 1 converging lens, 2 wave guide (UnitCube.cpp:91-157).
// With a prefix, the velocity, the right-hand sides and the solution are
// written as raw native-endian doubles (grid layout, complex interleaved) to
// <prefix>_velocity.bin, <prefix>_rhs.bin and <prefix>_solution.bin, and the
// object is Finalized and Initialized once more before the solve (the
// reference's Finalize keeps the object usable).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "psp.hpp"

using namespace psp;

int main(int argc, char* argv[])
{
    typedef Complex<double> C;
    psp::Initialize(argc, argv);
    const int n = argc > 1 ? std::atoi(argv[1]) : 40;
    const double omega = argc > 2 ? std::atof(argv[2]) : 30.;
    const int model = argc > 3 ? std::atoi(argv[3]) : 2;
    const bool fp32 = argc > 4 && std::atoi(argv[4]) != 0;
    const char* prefix = argc > 5 ? argv[5] : nullptr;
    auto dump = [&](const char* what, const void* data, size_t bytes) {
        const std::string name = std::string(prefix) + "_" + what + ".bin";
        FILE* f = std::fopen(name.c_str(), "wb");
        if (!f || std::fwrite(data, 1, bytes, f) != bytes) {
            std::fprintf(stderr, "could not write %s\n", name.c_str());
            std::exit(3);
        }
        std::fclose(f);
    };
    const int pmlSize = 5;
    const double sigma = 1.5, damping = 7.;
    const int planesPerPanel = 4;
    try {
        Discretization<double> disc(omega, n, n, n, 1., 1., 1., PML, PML, PML, PML, PML, pmlSize,
                                    sigma);
        DistHelmholtz<double> helmholtz(disc, mpi::COMM_WORLD, damping, planesPerPanel, 12,
                                        fp32 ? FRONTS_FP32 : FRONTS_FP64);
        DistUniformGrid<double> velocity(n, n, n, mpi::COMM_WORLD);
        double* v = velocity.Buffer();
        for (int z = 0; z < n; ++z)
            for (int y = 0; y < n; ++y)
                for (int x = 0; x < n; ++x) {
                    const double X = x / (n + 1.0), Y = y / (n + 1.0), Z = z / (n + 1.0);
                    const double aX = (X - 0.5) * (X - 0.5), aY = (Y - 0.5) * (Y - 0.5),
                                 aZ = (Z - 0.5) * (Z - 0.5);
                    double speed = 1.;
                    if (model == 1) speed = (1.0 - 0.4 * std::exp(-32. * (aX + aY + aZ))) / 0.8;
                    if (model == 2) speed = (1.0 - 0.4 * std::exp(-32. * (aX + aY))) / 0.8;
                    v[velocity.LocalIndex(x, y, z)] = speed;
                }
        auto t0 = std::chrono::steady_clock::now();
        helmholtz.Initialize(velocity);
        if (prefix) {
            helmholtz.Finalize();
            helmholtz.Initialize(velocity);
        }
        const double tInit =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

        DistUniformGrid<C> B(n, n, n, mpi::COMM_WORLD, 4);
        C* b = B.Buffer();
        const double c0[] = {0.5, 0.5, 0.1}, c1[] = {0.25, 0.25, 0.1}, c2[] = {0.75, 0.75, 0.5};
        const double dir[] = {0.57735, 0.57735, -0.57735};
        for (int z = 0; z < n; ++z)
            for (int y = 0; y < n; ++y)
                for (int x = 0; x < n; ++x) {
                    const double X = x / (n + 1.0), Y = y / (n + 1.0), Z = z / (n + 1.0);
                    auto sq = [](double a) { return a * a; };
                    const double a0 = sq(X - c0[0]) + sq(Y - c0[1]) + sq(Z - c0[2]);
                    const double a1 = sq(X - c1[0]) + sq(Y - c1[1]) + sq(Z - c1[2]);
                    const double a2 = sq(X - c2[0]) + sq(Y - c2[1]) + sq(Z - c2[2]);
                    const C f0 = n * std::exp(-10. * n * a0);
                    const C f1 = n * std::exp(-10. * n * a1);
                    const C f2 = n * std::exp(-10. * n * a2);
                    const C plane = std::exp(C(0, omega * (X * dir[0] + Y * dir[1] + Z * dir[2])));
                    const C beam = std::exp(-4. * omega * a2) * plane;
                    C* p = b + B.LocalIndex(x, y, z);
                    p[0] = -f0;
                    p[1] = -(f0 + f1 + f2);
                    p[2] = -beam;
                    const bool inner = x >= pmlSize && x < n - pmlSize && y >= pmlSize &&
                                       y < n - pmlSize && z >= pmlSize && z < n - pmlSize;
                    p[3] = inner ? -plane : C(0);
                }
        const size_t pts = (size_t)n * n * n;
        if (prefix) {
            dump("velocity", v, sizeof(double) * pts);
            dump("rhs", b, sizeof(C) * 4 * pts);
        }
        t0 = std::chrono::steady_clock::now();
        helmholtz.Solve(B, 20, 1e-5);
        const double tSolve =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        const psp_hip_solve_report& r = helmholtz.SolveReport();
        if (prefix) dump("solution", b, sizeof(C) * 4 * pts);
        std::printf("unit_cube n=%d omega=%g model=%d fronts=%s: initialize %.3f s, solve %.3f s, "
                    "%d GMRES iterations, max rel residual %.3e\n",
                    n, omega, model, fp32 ? "fp32" : "fp64", tInit, tSolve, r.iterations,
                    r.final_max_rel);
        helmholtz.Finalize();
        psp::Finalize();
        return r.final_max_rel < 1e-5 ? 0 : 1;
    } catch (std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        psp::Finalize();
        return 2;
    }
}
