// DistUniformGrid<F> at one process (include/psp/dist_uniform_grid.hpp:48-2198
// of the reference): numScalars values per point, points stored in the
// grid's GridDataOrder (XYZ: x + y*nx + z*nx*ny).  The on-disk formats
// (SequentialLoad, InterpolateTo, WritePlane, WriteVolume) run in the
// library's host code (psp_hip_grid_* in psp_hip.h).  GridData is an alias.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <string>
#include <type_traits>
#include <vector>

namespace psp {

enum GridDataOrder {
    XYZ = PSP_HIP_XYZ,
    XZY = PSP_HIP_XZY,
    YXZ = PSP_HIP_YXZ,
    YZX = PSP_HIP_YZX,
    ZXY = PSP_HIP_ZXY,
    ZYX = PSP_HIP_ZYX
};

enum PlaneType { XY = PSP_HIP_PLANE_XY, XZ = PSP_HIP_PLANE_XZ, YZ = PSP_HIP_PLANE_YZ };

template <typename F>
class DistUniformGrid {
    static_assert(std::is_same<F, double>::value || std::is_same<F, std::complex<double>>::value,
                  "grids hold double or std::complex<double>");
    static constexpr int kType =
        std::is_same<F, double>::value ? PSP_HIP_REAL64 : PSP_HIP_COMPLEX128;

public:
    DistUniformGrid(int nx, int ny, int nz, int /*px*/, int /*py*/, int /*pz*/, mpi::Comm comm,
                    int numScalars = 1, GridDataOrder order = XYZ)
        : DistUniformGrid(nx, ny, nz, comm, numScalars, order)
    {
    }
    DistUniformGrid(int nx, int ny, int nz, mpi::Comm comm = mpi::COMM_WORLD, int numScalars = 1,
                    GridDataOrder order = XYZ)
        : nx_(nx), ny_(ny), nz_(nz), numScalars_(numScalars), order_(order), comm_(comm),
          data_((size_t)nx * ny * nz * numScalars)
    {
        if (nx < 0 || ny < 0 || nz < 0) LogicError("grid dimensions must be non-negative");
    }

    int XSize() const { return nx_; }
    int YSize() const { return ny_; }
    int ZSize() const { return nz_; }
    int NumScalars() const { return numScalars_; }
    GridDataOrder Order() const { return order_; }
    mpi::Comm Comm() const { return comm_; }
    int XLocalSize() const { return nx_; }
    int YLocalSize() const { return ny_; }
    int ZLocalSize() const { return nz_; }
    int XShift() const { return 0; }
    int YShift() const { return 0; }
    int ZShift() const { return 0; }
    int XStride() const { return 1; }
    int YStride() const { return 1; }
    int ZStride() const { return 1; }
    int OwningProcess(int) const { return 0; }
    int OwningProcess(int, int, int) const { return 0; }
    int NaturalIndex(int x, int y, int z) const { return x + y * nx_ + z * nx_ * ny_; }
    int LocalIndex(int naturalIndex) const
    {
        return LocalIndex(naturalIndex % nx_, (naturalIndex / nx_) % ny_,
                          naturalIndex / (nx_ * ny_));
    }
    int LocalIndex(int x, int y, int z) const
    {
        const long long i = psp_hip_grid_index(nx_, ny_, nz_, order_, x, y, z);
        if (i < 0) LogicError("point (", x, ", ", y, ", ", z, ") is outside the grid");
        return (int)i * numScalars_;
    }
    F* Buffer() { return data_.data(); }
    const F* LockedBuffer() const { return data_.data(); }

    // Load the data from a raw binary file holding the points in this
    // grid's order (dist_uniform_grid.hpp:349-482).
    void SequentialLoad(std::string filename)
    {
        Check(psp_hip_grid_load(filename.c_str(), nx_, ny_, nz_, numScalars_, kType, order_,
                                data_.data()));
    }

    // Trilinear resampling onto nx x ny x nz points spanning the same box
    // (dist_uniform_grid.hpp:484-714).
    void InterpolateTo(int nx, int ny, int nz)
    {
        if (nx == nx_ && ny == ny_ && nz == nz_) return;
        std::vector<F> out((size_t)nx * ny * nz * numScalars_);
        Check(psp_hip_grid_interpolate(data_.data(), nx_, ny_, nz_, out.data(), nx, ny, nz,
                                       numScalars_, kType, order_));
        nx_ = nx;
        ny_ = ny;
        nz_ = nz;
        data_.swap(out);
    }

    void WritePlane(PlaneType planeType, int whichPlane, std::string basename) const
    {
        Check(psp_hip_grid_write_plane(data_.data(), nx_, ny_, nz_, numScalars_, kType, order_,
                                       planeType, whichPlane, basename.c_str()));
    }

    void WriteVolume(std::string basename) const
    {
        Check(psp_hip_grid_write_volume(data_.data(), nx_, ny_, nz_, numScalars_, kType, order_,
                                        basename.c_str()));
    }

    // The data in natural XYZ order, as the solver takes it.
    std::vector<F> Natural() const
    {
        if (order_ == XYZ) return data_;
        std::vector<F> out(data_.size());
        Check(psp_hip_grid_reorder(data_.data(), order_, out.data(), XYZ, nx_, ny_, nz_,
                                   numScalars_, kType));
        return out;
    }
    // Stores natural XYZ-order data in this grid's order.
    void SetNatural(const F* nat)
    {
        if (order_ == XYZ) {
            std::copy(nat, nat + data_.size(), data_.begin());
            return;
        }
        Check(psp_hip_grid_reorder(nat, XYZ, data_.data(), order_, nx_, ny_, nz_, numScalars_,
                                   kType));
    }

    static int FindCubicFactor(int commSize) { return FindFactor(commSize, 1. / 3.); }
    static int FindQuadraticFactor(int commSize) { return FindFactor(commSize, 1. / 2.); }

private:
    int nx_, ny_, nz_, numScalars_;
    GridDataOrder order_;
    mpi::Comm comm_;
    std::vector<F> data_;

    static void Check(int status)
    {
        if (status != PSP_HIP_OK) LogicError(psp_hip_grid_last_error());
    }
    // Smallest divisor of commSize at or above floor(commSize^p)
    // (dist_uniform_grid.hpp:176-193).
    static int FindFactor(int commSize, double p)
    {
        int f = std::max(1, (int)std::floor(std::pow((double)commSize, p)));
        while (commSize % f != 0) ++f;
        return f;
    }
};

template <typename F>
using GridData = DistUniformGrid<F>;

}  // namespace psp
