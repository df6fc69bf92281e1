"""vfem_dot_noalloc, the allocation-free variant of vfem_dot: device inputs, a
fixed per-device scratch array instead of library allocations.  Checked
against vfem_dot bit for bit (same kernels, same grid) on two calls in a row,
the second with a smaller input, through the C-ABI.  The input buffers are
allocated by the test with the HIP runtime libvfem.so itself is linked against."""
import ctypes

import numpy as np
import pytest

import vulkanfem_amd as V
from vulkanfem_amd import api, problems

pytestmark = pytest.mark.gpu

H2D = 1  # hipMemcpyHostToDevice


def _hip():
    V.lib()  # loads libvfem.so and with it its HIP runtime; the same handle is returned below
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


def _to_device(hip, a):
    p = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
    assert hip.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D) == 0
    return p


def _system(nx, ny, nz):
    p = problems.cantilever_box(nx, ny, nz)
    return V.assemble_system_matrix(p.voxels, problems.ke0(1.0, 0.3), p.fixed, 2)


def test_dot_noalloc_matches_dot_on_consecutive_calls():
    hip = _hip()
    rng = np.random.default_rng(7)
    big, small = _system(48, 24, 24), _system(12, 6, 6)
    assert small.n < big.n
    bufs = []
    for A in (big, small, big):
        a, b = rng.standard_normal(A.n), rng.standard_normal(A.n)
        ref = A.dot(a, b)
        da, db = _to_device(hip, a), _to_device(hip, b)
        bufs += [da, db]
        got = api.dot_noalloc(da.value, db.value, A.n)
        print(f"[no-alloc] n={A.n} dot={got!r} ref={ref!r}", flush=True)
        assert got == ref                       # same kernels and grid: bit for bit
        assert abs(got - a @ b) <= 1e-12 * np.abs(a * b).sum()
    # an empty input reduces to zero; a negative size is an argument error
    assert api.dot_noalloc(da.value, db.value, 0) == 0.0
    with pytest.raises(V.VfemError):
        api.dot_noalloc(da.value, db.value, -1)
    for p in bufs:
        assert hip.hipFree(p) == 0
    big.close()
    small.close()
