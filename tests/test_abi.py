"""C-ABI library checks that need no GPU: it loads, exports every symbol
include/psp_hip.h declares, its symbolic analysis equals the oracle's, and
its argument/state errors follow the reference's LogicError/RuntimeError
behaviour (no compute calls here)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle.oracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    from psp_amd import _lib
    L = _lib.lib()
    header = open(os.path.join(ROOT, "include", "psp_hip.h")).read()
    declared = set(re.findall(r"\b(psp_hip_\w+)\s*\(", header))
    assert declared, "no declarations parsed"
    assert declared == set(_lib.EXPORTS)
    for name in declared:
        assert hasattr(L, name), name


def test_no_device_reports_zero_devices_or_more():
    from psp_amd import _lib
    assert _lib.lib().psp_hip_device_count() >= 0


GRIDS = [(16, 16, 24, 12), (13, 9, 30, 3), (40, 40, 40, 12), (21, 17, 19, 5), (100, 100, 100, 12)]


@pytest.mark.parametrize("nx,ny,nz,cut", GRIDS)
def test_symbolic_analysis_equals_oracle(nx, ny, nz, cut):
    import psp_amd
    d = psp_amd.unit_cube_discretization(nx, nz=nz, ny=ny)
    h = psp_amd.DistHelmholtz(d, cutoff=cut)
    o = Oracle(d.omega, nx, ny, nz, d.wx, d.wy, d.wz, cutoff=cut)
    assert h.num_panels == o.num_panels
    for p in range(h.num_panels):
        assert h.panel_info(p) == o.panel_info(p)
    for c in range(3):
        if o.class_vsize(c) == 0:
            continue
        a, b = h.class_nodes(c), o.class_nodes(c)
        for key in ("off", "size", "parent", "usize"):
            assert np.array_equal(a[key], b[key]), key
        for s1, s2 in zip(a["structs"], b["structs"]):
            assert np.array_equal(s1, s2)
    for (x, y, v, vs) in ((0, 0, 0, 9), (nx - 1, ny - 1, 8, 9), (nx // 2, ny // 3, 4, 9)):
        assert h.reordered_index(x, y, v, vs) == o.reordered_index(x, y, v, vs)


def test_config_sizes_match_survey_table():
    """Fronts per panel and panel counts of the benchmark configurations."""
    import psp_amd
    for n, panels, fronts in ((40, 8, 255), (100, 23, 1293), (80, 18, 1023)):
        h = psp_amd.DistHelmholtz(psp_amd.unit_cube_discretization(n))
        assert h.num_panels == panels
        assert len(h.class_nodes(0)["off"]) == fronts


def test_shallow_domain_is_logic_error():
    import psp_amd
    with pytest.raises(psp_amd.LogicError):
        psp_amd.DistHelmholtz(psp_amd.unit_cube_discretization(10, nz=18))


def test_zero_frequency_is_logic_error():
    import psp_amd
    d = psp_amd.unit_cube_discretization(10, nz=30)
    d.omega = 0.0
    h = psp_amd.DistHelmholtz(d)
    with pytest.raises(psp_amd.LogicError, match="zero frequency"):
        h.Initialize(np.ones(10 * 10 * 30))


def test_velocity_size_mismatch_is_logic_error():
    import psp_amd
    h = psp_amd.DistHelmholtz(psp_amd.unit_cube_discretization(10, nz=30))
    with pytest.raises(psp_amd.LogicError, match="Velocity"):
        h.Initialize(np.ones(17))
    g = psp_amd.DistUniformGrid(2, 10, 10, 30, dtype=np.float64)
    with pytest.raises(psp_amd.LogicError, match="one entry per point"):
        h.Initialize(g)


def test_solve_before_initialize_is_logic_error():
    import psp_amd
    h = psp_amd.DistHelmholtz(psp_amd.unit_cube_discretization(10, nz=30))
    B = np.zeros((10 * 10 * 30, 1), dtype=np.complex128)
    with pytest.raises(psp_amd.LogicError):
        h.Solve(B)


def test_invalid_arguments_rejected_by_c_abi():
    from psp_amd import _lib
    L = _lib.lib()
    assert L.psp_hip_create(None, None) == _lib.ERR_INVALID
    out = C.c_void_p()
    bad = _lib.Desc(1.0, 0, 5, 5, 1, 1, 1, 0, 0, 0, 0, 0, 5, 5, 5, 1, 1, 1, 7.0, 4, 12, 0)
    assert L.psp_hip_create(C.byref(bad), C.byref(out)) == _lib.ERR_INVALID
    bad.nx = 5
    bad.precision = 7
    assert L.psp_hip_create(C.byref(bad), C.byref(out)) == _lib.ERR_INVALID
    assert L.psp_hip_last_error(None) == b"null context"


def test_python_api_mirrors_reference_names():
    import psp_amd
    for name in ("Initialize", "Solve", "SolveWithGMRES", "Finalize"):
        assert hasattr(psp_amd.DistHelmholtz, name)
    assert psp_amd.GridData is psp_amd.DistUniformGrid
    d = psp_amd.Discretization(1.0, 4, 4, 4, 1.0, 2.0, 1.0)
    assert d.sigmax == pytest.approx(2.4)  # 1.2 * max width (discretization.hpp:60-63)
    assert (d.bx, d.by, d.bz) == (5, 5, 5)


def test_solve_tiling_options_checked_on_host():
    """The front-solve tiling options (default shapes, persistent workgroups
    and their grid cap) rebuild the tile lists on the host before any device
    work, and out-of-range values are rejected with the library's message."""
    import psp_amd
    from psp_amd import workloads
    disc = workloads.discretization(12, 10, 24, 1.0, 10.0)
    h = psp_amd.DistHelmholtz(disc, cutoff=5)
    h.set_solve_tiling(persist=False)
    h.set_solve_tiling(persist=True, persist_grid=4, min_tiles=64, max_split=2, wide_wg=True)
    h.set_solve_tiling(persist_grid=0)
    with pytest.raises(psp_amd.LogicError, match="persistent grid cap"):
        h.set_solve_tiling(persist_grid=-1)
    with pytest.raises(psp_amd.LogicError, match="max split"):
        h.set_solve_tiling(max_split=0)
    with pytest.raises(psp_amd.LogicError, match="min tiles"):
        h.set_solve_tiling(min_tiles=0)


def test_finalize_keeps_the_object_usable_on_host():
    """Finalize frees the factors only (initialize_finalize.hpp:379-396): the
    object stays valid, Solve afterwards is a LogicError (not initialized),
    and Finalize may be called again; destroy releases the context."""
    import psp_amd
    h = psp_amd.DistHelmholtz(psp_amd.unit_cube_discretization(10, nz=30))
    panels = h.num_panels
    h.Finalize()
    h.Finalize()
    assert h.num_panels == panels
    B = np.zeros((10 * 10 * 30, 1), dtype=np.complex128)
    with pytest.raises(psp_amd.LogicError, match="before Initialize"):
        h.Solve(B)
    h.destroy()
    with pytest.raises(psp_amd.LogicError, match="destroyed"):
        h.Finalize()
