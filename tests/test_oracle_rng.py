"""The oracle's random-number generators (accesories.f90:38-147) and the order in
which the flow initialisation and the forcing consume them.

Pinned by: the Park-Miller "minimal standard" known answer (from seed 1 the
10,000th state is 1043618065, Park & Miller, CACM 31, 1988), and, where the
reference's own accesories.f90 has been compiled into oracle/_ref, bit-for-bit
agreement of ranff / ranpp / ranff2 / hpsort with it."""
import ctypes as C

import numpy as np
import pytest


def _ref():
    from oracle import oracle as O
    if not O.REF_LIB.exists():
        pytest.skip("oracle/_ref not built (reference sources absent)")
    lib = C.CDLL(str(O.REF_LIB))
    for name in ("ranff_", "ranpp_", "ranff2_"):
        getattr(lib, name).restype = C.c_double
    lib.hpsort_.restype = None
    return lib


def _state():
    return C.c_int(0), C.c_int(0), (C.c_int * 32)()


def test_park_miller_known_answer(oracle_mod):
    """ranff starts its table with 8 + 32 LCG steps and advances one step per draw,
    so after 1 + 9959 draws from seed 1 the state is x_10000 of 16807 x mod 2^31-1."""
    lib = oracle_mod.lib()
    idum, iy, iv = _state()
    idum.value = -1
    for _ in range(9960):
        lib.orc_ranff(C.byref(idum), C.byref(iy), iv)
    assert idum.value == 1043618065


def test_ranff_uniform_moments(oracle_mod):
    lib = oracle_mod.lib()
    idum, iy, iv = _state()
    idum.value = -232300
    x = np.array([lib.orc_ranff(C.byref(idum), C.byref(iy), iv) for _ in range(50000)])
    assert 0.0 < x.min() and x.max() < 1.0
    assert abs(x.mean() - 0.5) < 4 * np.sqrt(1 / 12 / x.size)
    assert abs(x.var() - 1 / 12) < 0.003


@pytest.mark.parametrize("seed", [-1, -123, -232300, -290710])
def test_ranff_matches_reference_build(oracle_mod, seed):
    ref = _ref()
    lib = oracle_mod.lib()
    a, b = _state(), _state()
    a[0].value = b[0].value = seed
    for _ in range(20000):
        x = lib.orc_ranff(C.byref(a[0]), C.byref(a[1]), a[2])
        y = ref.ranff_(C.byref(b[0]), C.byref(b[1]), b[2])
        assert x == y
    assert a[0].value == b[0].value and a[1].value == b[1].value and list(a[2]) == list(b[2])


@pytest.mark.parametrize("seed", [-7, -290711])
def test_ranpp_matches_reference_build(oracle_mod, seed):
    ref = _ref()
    lib = oracle_mod.lib()
    a, b = _state(), _state()
    a[0].value = b[0].value = seed
    for _ in range(20000):
        assert lib.orc_ranpp(C.byref(a[0]), C.byref(a[1]), a[2]) == ref.ranpp_(C.byref(b[0]), C.byref(b[1]), b[2])


@pytest.mark.parametrize("seed", [0, -10, -30])
def test_ranff2_matches_reference_build(oracle_mod, seed):
    """ranff2 (L'Ecuyer with shuffle) drives RELOCATE / OVERLAP_DET (seed -10 id)."""
    ref = _ref()
    lib = oracle_mod.lib()

    def st():
        return C.c_int(seed), (C.c_int * 32)(), C.c_int(0), C.c_int(0)

    a, b = st(), st()
    for _ in range(20000):
        x = lib.orc_ranff2(C.byref(a[0]), a[1], C.byref(a[2]), C.byref(a[3]))
        y = ref.ranff2_(C.byref(b[0]), b[1], C.byref(b[2]), C.byref(b[3]))
        assert x == y


def test_hpsort_reference_is_a_sort():
    """The removal list is sorted with hpsort (partlib.f90:1631); the oracle sorts
    with the C library.  Both must give the ascending order, duplicates kept."""
    ref = _ref()
    rng = np.random.default_rng(3)
    for n in (1, 2, 7, 100, 1001):
        a = rng.integers(1, 50, size=n).astype(np.int32)
        b = a.copy()
        ref.hpsort_(b.ctypes.data_as(C.POINTER(C.c_int)), C.byref(C.c_int(n)))
        assert np.array_equal(b, np.sort(a))
