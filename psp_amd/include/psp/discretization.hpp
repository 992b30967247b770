// Discretization<Real> (include/psp/discretization.hpp:15-110 of the
// reference): same fields, same four constructors, same defaults.
#pragma once

#include <algorithm>

namespace psp {

enum Boundary { PML = PSP_HIP_PML, DIRICHLET = PSP_HIP_DIRICHLET };
enum Stencil { SEVEN_POINT };

template <typename Real>
struct Discretization {
    Stencil stencil;
    Real omega;
    int nx, ny, nz;
    Real wx, wy, wz;
    Boundary frontBC, rightBC, backBC, leftBC, topBC;
    int bx, by, bz;
    Real sigmax, sigmay, sigmaz;

    Discretization(Real frequency, int xSize, int ySize, int zSize, Real xWidth, Real yWidth,
                   Real zWidth)
        : Discretization(frequency, xSize, ySize, zSize, xWidth, yWidth, zWidth, PML, PML, PML,
                         PML, PML, 5, 5, 5)
    {
    }

    Discretization(Real frequency, int xSize, int ySize, int zSize, Real xWidth, Real yWidth,
                   Real zWidth, Boundary front, Boundary right, Boundary back, Boundary left,
                   Boundary top)
        : Discretization(frequency, xSize, ySize, zSize, xWidth, yWidth, zWidth, front, right,
                         back, left, top, 5, 5, 5)
    {
    }

    Discretization(Real frequency, int xSize, int ySize, int zSize, Real xWidth, Real yWidth,
                   Real zWidth, Boundary front, Boundary right, Boundary back, Boundary left,
                   Boundary top, int xPMLSize, int yPMLSize, int zPMLSize)
        : stencil(SEVEN_POINT), omega(frequency), nx(xSize), ny(ySize), nz(zSize), wx(xWidth),
          wy(yWidth), wz(zWidth), frontBC(front), rightBC(right), backBC(back), leftBC(left),
          topBC(top), bx(xPMLSize), by(yPMLSize), bz(zPMLSize)
    {
        const Real maxDim = std::max(wx, std::max(wy, wz));
        sigmax = sigmay = sigmaz = Real(1.2) * maxDim;
    }

    Discretization(Real frequency, int xSize, int ySize, int zSize, Real xWidth, Real yWidth,
                   Real zWidth, Boundary front, Boundary right, Boundary back, Boundary left,
                   Boundary top, int pmlSize, double sigma)
        : stencil(SEVEN_POINT), omega(frequency), nx(xSize), ny(ySize), nz(zSize), wx(xWidth),
          wy(yWidth), wz(zWidth), frontBC(front), rightBC(right), backBC(back), leftBC(left),
          topBC(top), bx(pmlSize), by(pmlSize), bz(pmlSize), sigmax(sigma), sigmay(sigma),
          sigmaz(sigma)
    {
    }
};

}  // namespace psp
