// DistUniformGrid<F> at one process (include/psp/dist_uniform_grid.hpp:48-145,
// 302-342 of the reference): point-major storage with numScalars values
// interleaved per point, natural order x + y*nx + z*nx*ny.  GridData is an
// alias (the name the reference's GridDataOrder enum survives under).
#pragma once

#include <vector>

namespace psp {

template <typename F>
class DistUniformGrid {
public:
    DistUniformGrid(int nx, int ny, int nz, mpi::Comm comm = mpi::COMM_WORLD, int numScalars = 1)
        : nx_(nx), ny_(ny), nz_(nz), numScalars_(numScalars), comm_(comm),
          data_((size_t)nx * ny * nz * numScalars)
    {
    }
    int XSize() const { return nx_; }
    int YSize() const { return ny_; }
    int ZSize() const { return nz_; }
    int NumScalars() const { return numScalars_; }
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
    int OwningProcess(int, int, int) const { return 0; }
    int LocalIndex(int x, int y, int z) const { return (x + y * nx_ + z * nx_ * ny_) * numScalars_; }
    F* Buffer() { return data_.data(); }
    const F* LockedBuffer() const { return data_.data(); }

private:
    int nx_, ny_, nz_, numScalars_;
    mpi::Comm comm_;
    std::vector<F> data_;
};

template <typename F>
using GridData = DistUniformGrid<F>;

}  // namespace psp
