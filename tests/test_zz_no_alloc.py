"""The allocation-free glob (gt_glob_noalloc) against gt_glob, through the C
ABI: the same six sums (src/diagnostics.f90:125-179), written to device
memory with the partial sums in the library's static array.  Two calls in a
row, the second on a smaller grid (another context on the same GPU, so the
array holds the first call's larger contents), then the first again, then
right after a step; and a field whose only current device copy is in the
reference layout."""
import ctypes as C
import os

import numpy as np
import pytest

from helpers import params_for, random_spectral

pytestmark = pytest.mark.gpu


def _hip_runtime():
    """The HIP runtime the library is bound to (already loaded in this
    process): device memory for the output comes from it."""
    from gturbo_amd import _lib
    _lib.load()
    with open(f"/proc/{os.getpid()}/maps") as f:
        paths = {ln.split()[-1] for ln in f if "libamdhip64" in ln}
    assert paths, "no HIP runtime loaded"
    R = C.CDLL(sorted(paths)[0])
    R.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    R.hipFree.argtypes = [C.c_void_p]
    R.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    R.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return R


def _glob_noalloc(s, z, R, dev):
    R.hipMemset(dev, 0xFF, 48)  # NaN bytes: the call must write all six
    rc = s.L.gt_glob_noalloc(s.h, z.ctypes.data, dev)
    assert rc == 0, s.L.gt_last_error()
    s.synchronize()
    out = np.zeros(6)
    assert R.hipMemcpy(out.ctypes.data, dev, 48, 2) == 0  # device to host
    return out


def test_glob_noalloc_matches_glob():
    from gturbo_amd.solver import Solver
    R = _hip_runtime()
    dev = C.c_void_p()
    assert R.hipMalloc(C.byref(dev), 48) == 0
    try:
        runs = []
        for n, alpv in ((1024, 1.0), (256, 2.0)):  # the second grid is smaller
            s = Solver(params_for(n, alpv=alpv))
            z = random_spectral(n, 60 + n)
            s.enter_data(z)
            s.trunc(z)  # a spectral entry point: the compact copy is current
            want = s.glob(z)
            got = _glob_noalloc(s, z, R, dev)
            print(n, got, want, flush=True)
            assert np.array_equal(got, want), (n, got, want)
            runs.append((s, z, want))
        s, z, want = runs[0]  # the larger grid again, after the smaller one
        assert np.array_equal(_glob_noalloc(s, z, R, dev), want)
        # right after a step, whose last stage the context still holds back
        work = [s.field() for _ in range(4)]
        for a in work:
            s.enter_data(a, copyin=False)
        s.step4(z, *work)
        got = _glob_noalloc(s, z, R, dev)
        assert np.array_equal(got, s.glob(z)), "glob_noalloc saw the state before the step's end"
        # only the reference layout current: the whole half-spectrum is read
        n = 512
        s = Solver(params_for(n, alpv=0.5))
        zr = random_spectral(n, 7)
        s.enter_data(zr)
        got = _glob_noalloc(s, zr, R, dev)
        want = s.glob(zr)  # converts to the compact layout, sums in another order
        assert np.allclose(got, want, rtol=1e-13, atol=0), (got, want)
        for s, _, _ in runs:
            s.close()
        s.close()
    finally:
        R.hipFree(dev)
