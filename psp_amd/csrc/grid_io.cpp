// On-disk formats of the grid container at one process: raw binary loads in
// the six axis orders, trilinear resampling, and VTK ImageData writers for
// planes and volumes (DistUniformGrid<F>, include/psp/dist_uniform_grid.hpp
// of the reference).  Host-only C++; the data is the caller's buffer in the
// grid's own storage order.
#include <cerrno>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include <sys/stat.h>
#include <sys/types.h>

#include "psp_hip.h"

namespace {

thread_local std::string g_gridErr;

int GridFail(int code, const std::string& m)
{
    g_gridErr = m;
    return code;
}

bool ValidOrder(int o) { return o >= PSP_HIP_XYZ && o <= PSP_HIP_ZYX; }

// Point index of (x, y, z) in a grid stored in `order`, the fastest axis
// first (LocalIndex, dist_uniform_grid.hpp:302-342, and the file offsets of
// SequentialLoad, :349-482).  The reference's ZXY storage stride is
// zSize*ySize (:330), which only equals the file's zSize*xSize when nx == ny;
// both use z + x*nz + y*nz*nx here.
inline int64_t PointIndex(int order, int64_t nx, int64_t ny, int64_t nz, int64_t x, int64_t y,
                          int64_t z)
{
    switch (order) {
    case PSP_HIP_XZY: return x + z * nx + y * nx * nz;
    case PSP_HIP_YXZ: return y + x * ny + z * ny * nx;
    case PSP_HIP_YZX: return y + z * ny + x * ny * nz;
    case PSP_HIP_ZXY: return z + x * nz + y * nz * nx;
    case PSP_HIP_ZYX: return z + y * nz + x * nz * ny;
    default: return x + y * nx + z * nx * ny;
    }
}

size_t ScalarBytes(int type) { return type == PSP_HIP_COMPLEX128 ? 16 : 8; }

int CheckGrid(int nx, int ny, int nz, int numScalars, int type, int order)
{
    if (nx < 1 || ny < 1 || nz < 1) return GridFail(PSP_HIP_ERR_INVALID, "grid sizes must be >= 1");
    if (numScalars < 1) return GridFail(PSP_HIP_ERR_INVALID, "numScalars must be >= 1");
    if (type != PSP_HIP_REAL64 && type != PSP_HIP_COMPLEX128)
        return GridFail(PSP_HIP_ERR_INVALID, "unknown scalar type");
    if (!ValidOrder(order)) return GridFail(PSP_HIP_ERR_INVALID, "unknown GridDataOrder");
    return PSP_HIP_OK;
}

// Trilinear resampling of one scalar type (InterpolateTo,
// dist_uniform_grid.hpp:484-714): new point i maps to old coordinate
// i * (nOld - 1) / (nNew - 1); the upper neighbour is clamped at the last
// old plane; interpolation runs along z, then y, then x.
template <typename F>
void Resample(const F* in, int nx0, int ny0, int nz0, F* out, int nx, int ny, int nz, int ns,
              int order)
{
    const double xr = (nx - 1.0) / (nx0 - 1), yr = (ny - 1.0) / (ny0 - 1),
                 zr = (nz - 1.0) / (nz0 - 1);
    auto axis = [](int i, double ratio, int nOld, int& lo, int& hi, double& t) {
        const double o = nOld == 1 ? 0.0 : i / ratio;
        lo = (int)std::floor(o);
        if (lo > nOld - 1) lo = nOld - 1;
        if (lo < 0) lo = 0;
        hi = lo != nOld - 1 ? lo + 1 : lo;
        t = o - std::floor(o);
    };
    for (int z = 0; z < nz; ++z) {
        int z0, z1;
        double tz;
        axis(z, zr, nz0, z0, z1, tz);
        for (int y = 0; y < ny; ++y) {
            int y0, y1;
            double ty;
            axis(y, yr, ny0, y0, y1, ty);
            for (int x = 0; x < nx; ++x) {
                int x0, x1;
                double tx;
                axis(x, xr, nx0, x0, x1, tx);
                const int xs[2] = {x0, x1}, ys[2] = {y0, y1}, zs[2] = {z0, z1};
                const int64_t dst = PointIndex(order, nx, ny, nz, x, y, z) * ns;
                for (int s = 0; s < ns; ++s) {
                    F v[2][2][2];
                    for (int i = 0; i < 2; ++i)
                        for (int j = 0; j < 2; ++j)
                            for (int l = 0; l < 2; ++l)
                                v[i][j][l] =
                                    in[PointIndex(order, nx0, ny0, nz0, xs[i], ys[j], zs[l]) * ns + s];
                    for (int i = 0; i < 2; ++i)
                        for (int j = 0; j < 2; ++j)
                            v[i][j][0] = (1. - tz) * v[i][j][0] + tz * v[i][j][1];
                    for (int i = 0; i < 2; ++i) v[i][0][0] = (1. - ty) * v[i][0][0] + ty * v[i][1][0];
                    out[dst + s] = (1. - tx) * v[0][0][0] + tx * v[1][0][0];
                }
            }
        }
    }
}

// VTK ascii value: tiny magnitudes written as 0, default stream formatting
// (6 significant digits).
inline double Clean(double v) { return std::fabs(v) < 1.0e-300 ? 0.0 : v; }

std::string ImageHeader(const std::string& whole, const std::string& origin, double h,
                        const std::string& piece)
{
    std::ostringstream os;
    os << "<?xml version=\"1.0\"?>\n"
       << "<VTKFile type=\"ImageData\" version=\"0.1\">\n"
       << " <ImageData WholeExtent=\"" << whole << "\" "
       << "Origin=\"" << origin << "\" "
       << "Spacing=\"" << h << " " << h << " " << h << "\">\n"
       << "  <Piece Extent=\"" << piece << "\">\n"
       << "    <CellData Scalars=\"cell_scalars\">\n"
       << "     <DataArray type=\"Float64\" Name=\"cell_scalars\" format=\"ascii\">\n";
    return os.str();
}

const char* kImageFooter =
    "    </DataArray>\n"
    "   </CellData>\n"
    "  </Piece>\n"
    " </ImageData>\n"
    "</VTKFile>\n";

// Component c of a stored scalar: the value itself (real grids) or its real
// (c = 0) or imaginary (c = 1) part.
inline double Part(const void* data, int type, int64_t idx, int c)
{
    if (type == PSP_HIP_COMPLEX128) return ((const double*)data)[2 * idx + c];
    return ((const double*)data)[idx];
}

}  // namespace

extern "C" {

const char* psp_hip_grid_last_error(void) { return g_gridErr.c_str(); }

long long psp_hip_grid_index(int nx, int ny, int nz, int order, int x, int y, int z)
{
    if (!ValidOrder(order) || x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) return -1;
    return PointIndex(order, nx, ny, nz, x, y, z);
}

int psp_hip_grid_load(const char* filename, int nx, int ny, int nz, int numScalars, int type,
                      int order, void* data)
{
    int st = CheckGrid(nx, ny, nz, numScalars, type, order);
    if (st) return st;
    if (!filename || !data) return GridFail(PSP_HIP_ERR_INVALID, "null file name or buffer");
    std::ifstream is(filename, std::ios::in | std::ios::binary);
    if (!is.is_open()) return GridFail(PSP_HIP_ERR_INVALID, std::string("Could not open ") + filename);
    // The file holds the points in the grid's own order (at one process the
    // file offset of a point equals its storage offset), numScalars values
    // per point.
    const size_t bytes = ScalarBytes(type) * numScalars * (size_t)nx * ny * nz;
    is.read((char*)data, (std::streamsize)bytes);
    if ((size_t)is.gcount() != bytes)
        return GridFail(PSP_HIP_ERR_INVALID, std::string("file ") + filename + " holds " +
                                                 std::to_string(is.gcount()) + " of the " +
                                                 std::to_string(bytes) + " bytes of the grid");
    return PSP_HIP_OK;
}

int psp_hip_grid_reorder(const void* in, int inOrder, void* out, int outOrder, int nx, int ny,
                         int nz, int numScalars, int type)
{
    int st = CheckGrid(nx, ny, nz, numScalars, type, inOrder);
    if (st) return st;
    if (!ValidOrder(outOrder)) return GridFail(PSP_HIP_ERR_INVALID, "unknown GridDataOrder");
    if (!in || !out || in == out) return GridFail(PSP_HIP_ERR_INVALID, "bad buffers");
    const size_t eb = ScalarBytes(type) * numScalars;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x)
                std::memcpy((char*)out + eb * PointIndex(outOrder, nx, ny, nz, x, y, z),
                            (const char*)in + eb * PointIndex(inOrder, nx, ny, nz, x, y, z), eb);
    return PSP_HIP_OK;
}

int psp_hip_grid_interpolate(const void* in, int nx0, int ny0, int nz0, void* out, int nx, int ny,
                             int nz, int numScalars, int type, int order)
{
    int st = CheckGrid(nx0, ny0, nz0, numScalars, type, order);
    if (st) return st;
    if ((st = CheckGrid(nx, ny, nz, numScalars, type, order))) return st;
    if (!in || !out) return GridFail(PSP_HIP_ERR_INVALID, "null buffer");
    // (a single old plane along an axis is copied to every new plane; a
    // single new plane has no ratio to the old grid)
    if ((nx == 1 && nx0 > 1) || (ny == 1 && ny0 > 1) || (nz == 1 && nz0 > 1))
        return GridFail(PSP_HIP_ERR_INVALID, "cannot interpolate to a single plane");
    if (type == PSP_HIP_COMPLEX128)
        Resample((const std::complex<double>*)in, nx0, ny0, nz0, (std::complex<double>*)out, nx,
                 ny, nz, numScalars, order);
    else
        Resample((const double*)in, nx0, ny0, nz0, (double*)out, nx, ny, nz, numScalars, order);
    return PSP_HIP_OK;
}

int psp_hip_grid_write_plane(const void* data, int nx, int ny, int nz, int numScalars, int type,
                             int order, int planeType, int whichPlane, const char* basename)
{
    int st = CheckGrid(nx, ny, nz, numScalars, type, order);
    if (st) return st;
    if (!data || !basename) return GridFail(PSP_HIP_ERR_INVALID, "null buffer or basename");
    int n1, n2, depth;
    if (planeType == PSP_HIP_PLANE_XY) {
        n1 = nx, n2 = ny, depth = nz;
    } else if (planeType == PSP_HIP_PLANE_XZ) {
        n1 = nx, n2 = nz, depth = ny;
    } else if (planeType == PSP_HIP_PLANE_YZ) {
        n1 = ny, n2 = nz, depth = nx;
    } else {
        return GridFail(PSP_HIP_ERR_INVALID, "unknown PlaneType");
    }
    if (whichPlane < 0 || whichPlane >= depth) return GridFail(PSP_HIP_ERR_INVALID, "Invalid plane");
    // (WritePlaneHelper, dist_uniform_grid.hpp:722-1629) spacing from the
    // plane's own larger side; the plane sits at h * whichPlane on its
    // normal axis and is one cell thick.
    const double h = 1. / (std::max(n1, n2) + 1.0);
    std::ostringstream whole, origin;
    if (planeType == PSP_HIP_PLANE_XY) {
        whole << "0 " << nx << " 0 " << ny << " 0 1";
        origin << "0 0 " << h * whichPlane;
    } else if (planeType == PSP_HIP_PLANE_XZ) {
        whole << "0 " << nx << " 0 1 0 " << nz;
        origin << "0 " << h * whichPlane << " 0";
    } else {
        whole << "0 1 0 " << ny << " 0 " << nz;
        origin << h * whichPlane << " 0 0";
    }
    const std::string head = ImageHeader(whole.str(), origin.str(), h, whole.str());
    const int parts = type == PSP_HIP_COMPLEX128 ? 2 : 1;
    for (int k = 0; k < numScalars; ++k)
        for (int c = 0; c < parts; ++c) {
            std::ostringstream name;
            name << basename << "_" << k;
            if (parts == 2) name << (c == 0 ? "_real" : "_imag");
            name << ".vti";
            std::ofstream f(name.str(), std::ios::out);
            if (!f.is_open()) return GridFail(PSP_HIP_ERR_INVALID, "Could not open " + name.str());
            f << head;
            for (int j = 0; j < n2; ++j) {
                for (int i = 0; i < n1; ++i) {
                    int x, y, z;
                    if (planeType == PSP_HIP_PLANE_XY)
                        x = i, y = j, z = whichPlane;
                    else if (planeType == PSP_HIP_PLANE_XZ)
                        x = i, y = whichPlane, z = j;
                    else
                        x = whichPlane, y = i, z = j;
                    const int64_t idx = PointIndex(order, nx, ny, nz, x, y, z) * numScalars + k;
                    f << Clean(Part(data, type, idx, c)) << " ";
                }
                f << "\n";
            }
            f << kImageFooter;
            if (!f.good()) return GridFail(PSP_HIP_ERR_INVALID, "write failed: " + name.str());
        }
    return PSP_HIP_OK;
}

int psp_hip_grid_write_volume(const void* data, int nx, int ny, int nz, int numScalars, int type,
                              int order, const char* basename)
{
    int st = CheckGrid(nx, ny, nz, numScalars, type, order);
    if (st) return st;
    if (!data || !basename) return GridFail(PSP_HIP_ERR_INVALID, "null buffer or basename");
    // (WriteVolumeHelper, dist_uniform_grid.hpp:1786-2198) at one process:
    // a parallel description basename_k[_part].pvti naming one piece, the
    // whole box, stored as basename_k[_part]/0.vti; spacing from the largest
    // side.
    const double h = 1.0 / (std::max(nx, std::max(ny, nz)) + 1.0);
    std::ostringstream ext;
    ext << "0 " << nx << " 0 " << ny << " 0 " << nz;
    const int parts = type == PSP_HIP_COMPLEX128 ? 2 : 1;
    for (int k = 0; k < numScalars; ++k)
        for (int c = 0; c < parts; ++c) {
            std::ostringstream stem;
            stem << basename << "_" << k;
            if (parts == 2) stem << (c == 0 ? "_real" : "_imag");
            {
                std::ofstream p(stem.str() + ".pvti", std::ios::out);
                if (!p.is_open())
                    return GridFail(PSP_HIP_ERR_INVALID, "Could not open " + stem.str() + ".pvti");
                p << "<?xml version=\"1.0\"?>\n"
                  << "<VTKFile type=\"PImageData\" version=\"0.1\">\n"
                  << " <PImageData WholeExtent=\"" << ext.str() << "\" "
                  << "Origin=\"0 0 0\" "
                  << "Spacing=\"" << h << " " << h << " " << h << "\" "
                  << "GhostLevel=\"0\">\n"
                  << "  <PCellData Scalars=\"cell_scalars\">\n"
                  << "    <PDataArray type=\"Float64\" Name=\"cell_scalars\"/>\n"
                  << "  </PCellData>\n"
                  << "  <Piece Extent=\"" << ext.str() << "\" Source=\"" << stem.str()
                  << "/0.vti\"/>\n"
                  << " </PImageData>\n"
                  << "</VTKFile>\n";
                if (!p.good()) return GridFail(PSP_HIP_ERR_INVALID, "write failed: " + stem.str());
            }
            if (mkdir(stem.str().c_str(), 0755) != 0 && errno != EEXIST)
                return GridFail(PSP_HIP_ERR_INVALID, "Could not create directory " + stem.str());
            const std::string piece = stem.str() + "/0.vti";
            std::ofstream f(piece, std::ios::out);
            if (!f.is_open()) return GridFail(PSP_HIP_ERR_INVALID, "Could not open " + piece);
            f << ImageHeader(ext.str(), "0 0 0", h, ext.str());
            for (int z = 0; z < nz; ++z)
                for (int y = 0; y < ny; ++y) {
                    for (int x = 0; x < nx; ++x) {
                        const int64_t idx = PointIndex(order, nx, ny, nz, x, y, z) * numScalars + k;
                        f << Clean(Part(data, type, idx, c)) << " ";
                    }
                    f << "\n";
                }
            f << kImageFooter;
            if (!f.good()) return GridFail(PSP_HIP_ERR_INVALID, "write failed: " + piece);
        }
    return PSP_HIP_OK;
}

}  // extern "C"
