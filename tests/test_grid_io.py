"""Grid container on-disk formats (host only): SequentialLoad in the six
GridDataOrders, trilinear InterpolateTo, VTK WritePlane / WriteVolume
(include/psp/dist_uniform_grid.hpp:349-2198 of the reference).

Expected files are built independently of the library: raw files by numpy
transposes of an (z, y, x) array, interpolation targets from closed forms
that trilinear interpolation reproduces exactly (or within its h^2 bound),
and the VTK text written out by hand from the reference's writers."""
import os
import subprocess

import numpy as np
import pytest

import psp_amd
from psp_amd import DistUniformGrid, LogicError

ORDERS = {
    # order: axes of the stored array from slowest to fastest, as (z, y, x) indices
    psp_amd.XYZ: (0, 1, 2),  # z, y, x  -> x fastest
    psp_amd.XZY: (1, 0, 2),  # y, z, x
    psp_amd.YXZ: (0, 2, 1),  # z, x, y
    psp_amd.YZX: (2, 0, 1),  # x, z, y
    psp_amd.ZXY: (1, 2, 0),  # y, x, z
    psp_amd.ZYX: (2, 1, 0),  # x, y, z
}


def field(nx, ny, nz, ns=1, cplx=False):
    """A[z, y, x, s] with every entry distinct."""
    z, y, x, s = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), np.arange(ns),
                             indexing="ij")
    a = 1.0 + x + 100.0 * y + 10000.0 * z + 0.25 * s
    if cplx:
        a = a + 1j * (7.0 - 3 * x + 11 * y - 0.5 * z + s)
    return a


@pytest.mark.parametrize("order", sorted(ORDERS))
@pytest.mark.parametrize("cplx", [False, True])
def test_sequential_load_every_order(tmp_path, order, cplx):
    nx, ny, nz, ns = 3, 4, 5, 2
    A = field(nx, ny, nz, ns, cplx)
    perm = ORDERS[order]
    raw = np.ascontiguousarray(A.transpose(*perm, 3))  # points in file order, ns per point
    path = tmp_path / "grid.dat"
    raw.astype(np.complex128 if cplx else np.float64).tofile(path)
    g = DistUniformGrid(ns, nx, ny, nz, dtype=np.complex128 if cplx else np.float64, order=order)
    g.SequentialLoad(str(path))
    flat = g.Buffer().reshape(-1)
    for (z, y, x) in [(0, 0, 0), (4, 3, 2), (1, 2, 0), (3, 0, 1), (2, 3, 2)]:
        li = g.LocalIndex(x, y, z)
        np.testing.assert_array_equal(flat[li:li + ns], A[z, y, x])
    nat = g.natural().reshape(nz, ny, nx, ns)
    np.testing.assert_array_equal(nat, A)


def test_sequential_load_missing_and_short_files(tmp_path):
    g = DistUniformGrid(1, 3, 3, 3, dtype=np.float64)
    with pytest.raises(LogicError, match="Could not open"):
        g.SequentialLoad(str(tmp_path / "absent.dat"))
    path = tmp_path / "short.dat"
    np.zeros(26).tofile(path)
    with pytest.raises(LogicError, match="bytes"):
        g.SequentialLoad(str(path))


def test_reorder_round_trip_and_local_index_bounds():
    nx, ny, nz = 4, 3, 2
    A = field(nx, ny, nz, 1, True)
    for order in ORDERS:
        g = DistUniformGrid(1, nx, ny, nz, order=order)
        g.set_natural(A.reshape(-1, 1))
        np.testing.assert_array_equal(g.natural().reshape(nz, ny, nx, 1), A)
        stored = g.Buffer().reshape(-1)
        np.testing.assert_array_equal(stored, A[..., 0].transpose(*ORDERS[order]).reshape(-1))
    with pytest.raises(LogicError):
        DistUniformGrid(1, nx, ny, nz).LocalIndex(nx, 0, 0)


def _coords(n):
    return np.arange(n) / (n - 1.0)


@pytest.mark.parametrize("order", [psp_amd.XYZ, psp_amd.ZYX, psp_amd.YXZ])
def test_interpolate_reproduces_trilinear_functions(order):
    """f = 1 + 2X - 3Y + Z/2 + XY - 4YZ + 5XYZ is trilinear: resampling from
    any grid spanning [0,1]^3 reproduces it exactly (up to rounding)."""
    def f(X, Y, Z):
        return 1 + 2 * X - 3 * Y + Z / 2 + X * Y - 4 * Y * Z + 5 * X * Y * Z + 0j

    src = (5, 4, 6)
    for dst in [(9, 7, 11), (3, 3, 4), (5, 4, 6), (13, 2, 8)]:
        Z, Y, X = np.meshgrid(_coords(src[2]), _coords(src[1]), _coords(src[0]), indexing="ij")
        g = DistUniformGrid(1, *src, order=order)
        g.set_natural(f(X, Y, Z).reshape(-1, 1))
        g.InterpolateTo(*dst)
        assert (g.XSize(), g.YSize(), g.ZSize()) == dst
        Z, Y, X = np.meshgrid(_coords(dst[2]), _coords(dst[1]), _coords(dst[0]), indexing="ij")
        np.testing.assert_allclose(g.natural().reshape(-1), f(X, Y, Z).reshape(-1), rtol=0,
                                   atol=1e-12)


def test_interpolate_quadratic_within_h2_bound_and_exact_on_old_points():
    """X^2 on a coarse grid: linear interpolation errs by at most h^2/4 at the
    midpoints (h^2/8 of the second derivative 2), and is exact where new
    points land on old ones (every other point when n_new = 2 n_old - 1)."""
    n0, n1 = 6, 11
    g = DistUniformGrid(1, n0, 2, 2, dtype=np.float64)
    X = _coords(n0)
    g.set_natural(np.tile(X ** 2, 4).reshape(-1, 1))
    g.InterpolateTo(n1, 2, 2)
    got = g.natural().reshape(2, 2, n1)[0, 0]
    Xn = _coords(n1)
    h = 1.0 / (n0 - 1)
    assert np.max(np.abs(got - Xn ** 2)) <= h * h / 4 + 1e-15
    np.testing.assert_allclose(got[::2], X ** 2, rtol=0, atol=1e-15)
    mid = got[1::2]
    np.testing.assert_allclose(mid, (X[:-1] ** 2 + X[1:] ** 2) / 2, rtol=0, atol=1e-15)


def test_interpolate_same_size_is_a_no_op():
    g = DistUniformGrid(2, 3, 4, 5)
    g.set_natural(field(3, 4, 5, 2, True).reshape(-1, 2))
    before = g.Buffer().copy()
    g.InterpolateTo(3, 4, 5)
    np.testing.assert_array_equal(g.Buffer(), before)


HEADER = ('<?xml version="1.0"?>\n'
          '<VTKFile type="ImageData" version="0.1">\n'
          ' <ImageData WholeExtent="{ext}" Origin="{origin}" Spacing="{h} {h} {h}">\n'
          '  <Piece Extent="{ext}">\n'
          '    <CellData Scalars="cell_scalars">\n'
          '     <DataArray type="Float64" Name="cell_scalars" format="ascii">\n')
FOOTER = ('    </DataArray>\n'
          '   </CellData>\n'
          '  </Piece>\n'
          ' </ImageData>\n'
          '</VTKFile>\n')


def test_write_plane_real_exact_text(tmp_path):
    # 2 x 3 x 2 grid, value x + 100 y + 10000 z (+ 0.25 for scalar 1)
    nx, ny, nz = 2, 3, 2
    A = field(nx, ny, nz, 2) - 1.0
    g = DistUniformGrid(2, nx, ny, nz, dtype=np.float64)
    g.set_natural(A.reshape(-1, 2))
    base = str(tmp_path / "p")
    g.WritePlane(psp_amd.XY, 1, base)
    h = 1.0 / 4.0  # max(nx, ny) + 1
    body0 = "".join(" ".join(f"{A[1, y, x, 0]:g}" for x in range(nx)) + " \n" for y in range(ny))
    want0 = HEADER.format(ext="0 2 0 3 0 1", origin=f"0 0 {h:g}", h=f"{h:g}") + body0 + FOOTER
    assert open(base + "_0.vti").read() == want0
    body1 = "".join(" ".join(f"{A[1, y, x, 1]:g}" for x in range(nx)) + " \n" for y in range(ny))
    assert body1 in open(base + "_1.vti").read()
    # XZ plane y = 2: rows are z, columns x; YZ plane x = 1: rows z, columns y
    g.WritePlane(psp_amd.XZ, 2, base + "xz")
    h2 = 1.0 / 3.0
    bodyxz = "".join(" ".join(f"{A[z, 2, x, 0]:g}" for x in range(nx)) + " \n" for z in range(nz))
    assert open(base + "xz_0.vti").read() == HEADER.format(
        ext="0 2 0 1 0 2", origin=f"0 {2 * h2:g} 0", h=f"{h2:g}") + bodyxz + FOOTER
    g.WritePlane(psp_amd.YZ, 1, base + "yz")
    h3 = 1.0 / 4.0
    bodyyz = "".join(" ".join(f"{A[z, y, 1, 0]:g}" for y in range(ny)) + " \n" for z in range(nz))
    assert open(base + "yz_0.vti").read() == HEADER.format(
        ext="0 1 0 3 0 2", origin=f"{h3:g} 0 0", h=f"{h3:g}") + bodyyz + FOOTER
    with pytest.raises(LogicError, match="Invalid plane"):
        g.WritePlane(psp_amd.XY, nz, base)


def test_write_plane_complex_splits_parts_and_zeroes_denormals(tmp_path):
    g = DistUniformGrid(1, 2, 2, 1)
    g.set_natural(np.array([[1.5 - 2j], [1e-310 + 0.1j], [-3 + 0j], [2.25 + 1e-301j]]))
    base = str(tmp_path / "c")
    g.WritePlane(psp_amd.XY, 0, base)
    re = open(base + "_0_real.vti").read().split('format="ascii">\n')[1].split("    </")[0]
    im = open(base + "_0_imag.vti").read().split('format="ascii">\n')[1].split("    </")[0]
    assert re == "1.5 0 \n-3 2.25 \n"
    assert im == "-2 0.1 \n0 0 \n"


def test_write_volume_layout(tmp_path):
    nx, ny, nz = 3, 2, 2
    A = field(nx, ny, nz, 1) * 0.5
    g = DistUniformGrid(1, nx, ny, nz, dtype=np.float64, order=psp_amd.ZXY)
    g.set_natural(A.reshape(-1, 1))
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        g.WriteVolume("vol")
    finally:
        os.chdir(cwd)
    h = 1.0 / 4.0
    pv = open(tmp_path / "vol_0.pvti").read()
    assert pv == ('<?xml version="1.0"?>\n'
                  '<VTKFile type="PImageData" version="0.1">\n'
                  f' <PImageData WholeExtent="0 3 0 2 0 2" Origin="0 0 0" Spacing="{h:g} {h:g} {h:g}" '
                  'GhostLevel="0">\n'
                  '  <PCellData Scalars="cell_scalars">\n'
                  '    <PDataArray type="Float64" Name="cell_scalars"/>\n'
                  '  </PCellData>\n'
                  '  <Piece Extent="0 3 0 2 0 2" Source="vol_0/0.vti"/>\n'
                  ' </PImageData>\n'
                  '</VTKFile>\n')
    body = "".join(" ".join(f"{A[z, y, x, 0]:g}" for x in range(nx)) + " \n"
                   for z in range(nz) for y in range(ny))
    want = HEADER.format(ext="0 3 0 2 0 2", origin="0 0 0", h=f"{h:g}") + body + FOOTER
    assert open(tmp_path / "vol_0" / "0.vti").read() == want


def test_write_volume_complex_writes_real_and_imag(tmp_path):
    g = DistUniformGrid(2, 2, 2, 2)
    g.set_natural(field(2, 2, 2, 2, True).reshape(-1, 2))
    base = str(tmp_path / "cv")
    g.WriteVolume(base)
    for k in (0, 1):
        for part in ("real", "imag"):
            assert os.path.exists(f"{base}_{k}_{part}.pvti")
            assert os.path.exists(f"{base}_{k}_{part}/0.vti")


CPP = r'''
#include <cstdio>
#include "psp.hpp"
using namespace psp;
int main(int argc, char** argv)
{
    psp::Initialize(argc, argv);
    DistUniformGrid<double> v(4, 3, 2, mpi::COMM_WORLD, 1, YZX);
    v.SequentialLoad(argv[1]);
    double s = 0;
    for (int z = 0; z < 2; ++z)
        for (int y = 0; y < 3; ++y)
            for (int x = 0; x < 4; ++x) s += v.Buffer()[v.LocalIndex(x, y, z)] * (1 + x + 4 * y + 12 * z);
    v.InterpolateTo(7, 5, 3);
    v.WritePlane(XZ, 4, argv[2]);
    std::printf("%.17g %d %d %d %d\n", s, v.XSize(), v.YSize(), v.ZSize(),
                DistUniformGrid<double>::FindCubicFactor(12));
    try {
        v.SequentialLoad("/nonexistent/file");
    } catch (std::logic_error& e) {
        std::printf("caught: %s\n", e.what());
    }
    return 0;
}
'''


def test_cpp_grid_api_on_host(tmp_path):
    """The C++ mirror of the grid class, compiled and run on the host."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = tmp_path / "g.cpp"
    src.write_text(CPP)
    exe = tmp_path / "g"
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{root}/psp_amd/include", f"-I{root}/include",
                    str(src), "-o", str(exe), f"-L{root}/psp_amd", "-lpsp_hip",
                    f"-Wl,-rpath,{root}/psp_amd"], check=True)
    A = field(4, 3, 2, 1)[..., 0]
    A.transpose(*ORDERS[psp_amd.YZX]).astype(np.float64).tofile(tmp_path / "v.dat")
    out = subprocess.run([str(exe), str(tmp_path / "v.dat"), str(tmp_path / "pl")], check=True,
                         capture_output=True, text=True).stdout.split("\n")
    z, y, x = np.meshgrid(np.arange(2), np.arange(3), np.arange(4), indexing="ij")
    want = float(np.sum(A * (1 + x + 4 * y + 12 * z)))
    s, nx, ny, nz, f = out[0].split()
    assert float(s) == want and (int(nx), int(ny), int(nz)) == (7, 5, 3) and int(f) == 2
    assert "caught: Could not open /nonexistent/file" in out[1]
    assert os.path.exists(tmp_path / "pl_0.vti")
