"""The allocation-free E(k) entry point (ddns_get_spectrum_device) against
ddns_get_spectrum on the same u-hat.  Both run the same two kernels in the same
fixed summation order, so the shell sums must be identical bit for bit.  Two calls in
a row share the per-GPU scratch array: the second, on a smaller grid (fewer shells),
must not see anything the first left behind; a third call repeats the first grid.
The output array is device memory from the HIP runtime that libddns.so itself uses
(libamdhip64.so.7 of /opt/rocm), filled with NaN first, so a shell the call fails to
write shows up."""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import random_spectral

pytestmark = pytest.mark.gpu


def _hip():
    hip = C.CDLL("libamdhip64.so.7")          # already loaded by libddns.so: the same runtime
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return hip


def _spectrum_device(ddns, s, nek):
    hip = _hip()
    H2D, D2H = 1, 2
    dev = C.c_void_p()
    assert hip.hipMalloc(C.byref(dev), 8 * nek) == 0
    try:
        host = np.full(nek, np.nan)
        assert hip.hipMemcpy(dev, host.ctypes.data, 8 * nek, H2D) == 0
        st = s.lib.ddns_get_spectrum_device(s.h, dev, None)
        assert st == 0, s.lib.ddns_last_error(s.h).decode()
        assert hip.hipDeviceSynchronize() == 0          # the handle's stream: wait for the device
        out = np.empty(nek)
        assert hip.hipMemcpy(out.ctypes.data, dev, 8 * nek, D2H) == 0
    finally:
        hip.hipFree(dev)
    return out


@pytest.mark.parametrize("prec_name", ["FP64", "FP32_MIXED"])
def test_spectrum_device_equals_spectrum(ddns, prec_name):
    prec = getattr(ddns, prec_name)
    for N, seed in ((64, 11), (32, 12), (64, 13)):
        p = ddns.default_params(N=N, precision=prec, npart=0)
        v = random_spectral(N, seed=seed, kmax=N // 3)
        with ddns.Solver(p) as s:
            s.set_flow(v, v)
            s.step(1)
            ref = s.spectrum()
            nek = s.lib.ddns_nek(s.h)
            got = _spectrum_device(ddns, s, nek)
        assert ref.shape == (nek,) and got.shape == (nek,)
        assert np.all(np.isfinite(got)), (N, seed)
        assert ref.max() > 0
        assert np.array_equal(got, ref), (N, seed, np.abs(got - ref).max())
