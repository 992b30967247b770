"""psp_hip_precondition_noalloc (the sweep with its buffers in the library's
fixed device scratch) against psp_hip_precondition, through the C ABI: two
calls in a row on different contexts, the second with a smaller grid and
fewer right-hand sides, each bitwise equal to the allocating entry point's
result (same kernels, same tile shapes, same summation order).  A request
beyond the scratch's size is an error, not a fault."""
import ctypes as C

import numpy as np
import pytest

from conftest import rand_block

pytestmark = pytest.mark.gpu


def _ctx_and_block(torch, nx, ny, nz, k, precision, seed):
    import psp_amd
    from psp_amd import workloads
    vel = workloads.bump_velocity(nx, ny, nz, cfg=seed)
    disc = workloads.discretization(nx, ny, nz, float(vel.min()), 10.0)
    h = psp_amd.DistHelmholtz(disc, precision=precision)
    h.Initialize(vel)
    X = rand_block(np.random.default_rng(seed), nx * ny * nz, k)
    return h, torch.from_numpy(X).cuda()


def _call(lib, name, h, B):
    return getattr(lib, name)(h._ctx(), C.c_void_p(B.data_ptr()), B.shape[1])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_precondition_equals_precondition(torch_cuda, precision):
    torch = torch_cuda
    from psp_amd import _lib
    lib = _lib.lib()
    cases = [(24, 22, 30, 8, 31), (16, 14, 26, 3, 32)]  # second call: smaller input
    for nx, ny, nz, k, seed in cases:
        h, X = _ctx_and_block(torch, nx, ny, nz, k, precision, seed)
        ref = X.clone()
        got = X.clone()
        assert _call(lib, "psp_hip_precondition", h, ref) == _lib.OK
        h._bind_stream()
        assert _call(lib, "psp_hip_precondition_noalloc", h, got) == _lib.OK
        torch.cuda.synchronize()
        assert torch.equal(got, ref), (nx, ny, nz, k)
        assert torch.isfinite(got).all()
        del h


def test_noalloc_precondition_rejects_oversized_requests(torch_cuda):
    torch = torch_cuda
    from psp_amd import _lib
    from psp_amd.helmholtz import LogicError
    h, _ = _ctx_and_block(torch, 24, 22, 30, 1, "fp64", 33)
    big = torch.zeros((24 * 22 * 30, 256), dtype=torch.complex128, device="cuda")
    with pytest.raises(LogicError, match="scratch"):
        h.PreconditionNoAlloc(big)
    assert _call(_lib.lib(), "psp_hip_precondition_noalloc", h, big) == 1  # PSP_HIP_ERR_INVALID
