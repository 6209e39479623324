"""Python wrapper of the oracle (liboracle.so, the FP64 CPU restatement).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.  The API mirrors the GPU Solver so a test can drive
both with the same calls.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from mpi_2d_domain_decomp_turb_particles_amd import _abi
from mpi_2d_domain_decomp_turb_particles_amd._abi import CollisionCounts, Monitors, NBIN, Params

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libref_acc.so"

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_P = C.c_void_p


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def load() -> C.CDLL:
    if not LIB.exists():
        build()
    lib = C.CDLL(str(LIB))
    sig = {
        "orc_create": (C.c_int, [C.POINTER(Params), C.POINTER(_P)]),
        "orc_destroy": (None, [_P]),
        "orc_nek": (C.c_int, [_P]),
        "orc_set_flow": (C.c_int, [_P] + [_D] * 6),
        "orc_init_flow": (C.c_int, [_P, C.c_int, C.c_int, C.c_double, C.c_double]),
        "orc_splitmix64": (C.c_uint64, [C.c_uint64, C.c_uint64]),
        "orc_get_flow": (C.c_int, [_P] + [_D] * 6),
        "orc_set_forcing_state": (C.c_int, [_P, _D, C.c_int, C.c_int, _I]),
        "orc_get_forcing_state": (C.c_int, [_P, _D, _I, _I, _I]),
        "orc_set_droplets": (C.c_int, [_P, C.c_int, _D, _D, _D, _D]),
        "orc_get_droplets": (C.c_int, [_P, _I, _D, _D, _D, _D, _D, _D]),
        "orc_set_new_particle": (C.c_int, [_P, C.c_int]),
        "orc_num_threads": (C.c_int, []),
        "orc_set_istep": (C.c_int, [_P, C.c_int]),
        "orc_step": (C.c_int, [_P, C.c_int]),
        "orc_flow_only": (C.c_int, [_P, C.c_int]),
        "orc_get_spectrum": (C.c_int, [_P, _D]),
        "orc_get_spectra": (C.c_int, [_P, _D, _D, _D]),
        "orc_get_monitors": (C.c_int, [_P, C.c_int, C.POINTER(Monitors)]),
        "orc_get_physical_velocity": (C.c_int, [_P, _D, _D, _D]),
        "orc_get_collision_counts": (C.c_int, [_P, C.POINTER(CollisionCounts)]),
        "orc_get_collision_bins": (C.c_int, [_P, _I] + [_D] * 7),
        "orc_get_collision_pairs": (C.c_int, [_P, C.POINTER(C.c_int64), C.c_int, _I]),
        "orc_get_removal_list": (C.c_int, [_P, _I, C.c_int, _I]),
        "orc_get_contact_records": (C.c_int, [_P, _D, C.c_int, _I]),
        "orc_collision_detect": (C.c_int, [_P, C.c_int] + [_D] * 5),
        "orc_overlap_fix": (C.c_int, [_P]),
        "orc_get_relocate_state": (C.c_int, [_P, _I, _I, _I, _I]),
        "orc_get_scaling": (C.c_int, [_P, _D]),
        "orc_fft_r2c": (C.c_int, [C.c_int, _D]),
        "orc_fft_c2r": (C.c_int, [C.c_int, _D]),
        "orc_fft1d": (C.c_int, [C.c_int, _D, C.c_int]),
        "orc_ranff": (C.c_double, [_I, _I, _I]),
        "orc_ranpp": (C.c_double, [_I, _I, _I]),
        "orc_ranff2": (C.c_double, [_I, _I, _I, _I]),
        "orc_interp_weights": (C.c_int, [C.c_double, C.c_double, _I, _D]),
        "orc_symmetrize": (C.c_int, [C.c_int, _D]),
        "orc_dealias": (C.c_int, [C.c_int, _D]),
        "orc_supfor_det": (C.c_int, [C.c_int, _D, _D, _D, _D, _D]),
        "orc_supfor_sth": (C.c_int, [_P, _D, _D, _D, _D]),
        "orc_shell_counts": (C.c_int, [C.c_int, _D]),
        "orc_hdi_solve": (C.c_int, [_P, C.c_int, _D, _D, _D, _D, _I, _D]),
        "orc_init_droplets": (C.c_int, [_P, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]),
    }
    for k, (r, a) in sig.items():
        f = getattr(lib, k)
        f.restype = r
        f.argtypes = a
    return lib


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        _lib = load()
    return _lib


def dp(a):
    if a is None:
        return None
    if a.dtype != np.float64 or not a.flags["F_CONTIGUOUS"]:
        # every multi-dimensional array crosses in Fortran (column-major) order
        raise TypeError("expected a Fortran-contiguous float64 array")
    return a.ctypes.data_as(_D)


def ip(a):
    return None if a is None else a.ctypes.data_as(_I)


def spectral_zeros(N):
    return np.zeros((N + 2, N + 2, N), order="F")


class Oracle:
    """FP64 CPU restatement; same method names as the GPU Solver."""

    def __init__(self, params: Params):
        self.lib = lib()
        self.p = params
        self.N = params.nx
        h = C.c_void_p()
        st = self.lib.orc_create(C.byref(params), C.byref(h))
        if st != 0:
            raise RuntimeError(f"orc_create failed: {st}")
        self.h = h
        self.nek = self.lib.orc_nek(h)

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, st):
        if st < 0:
            raise RuntimeError(f"oracle error {st}")
        return st

    def init_flow(self, kind, seed, k0=4.0, urms=1.0):
        self._chk(self.lib.orc_init_flow(self.h, kind, seed, k0, urms))

    def set_flow(self, v, o=None):
        v = [np.asfortranarray(x, dtype=np.float64) for x in v]
        o = [np.asfortranarray(x, dtype=np.float64) for x in o] if o is not None else [None] * 3
        self._chk(self.lib.orc_set_flow(self.h, *[dp(x) for x in v], *[dp(x) for x in o]))

    def get_flow(self):
        a = [spectral_zeros(self.N) for _ in range(6)]
        self._chk(self.lib.orc_get_flow(self.h, *[dp(x) for x in a]))
        return a[:3], a[3:]

    def set_forcing_state(self, b, iseedf, iyf, ivf):
        b = np.ascontiguousarray(b, dtype=np.float64)
        ivf = np.ascontiguousarray(ivf, dtype=np.int32)
        self._chk(self.lib.orc_set_forcing_state(self.h, dp(b), iseedf, iyf, ip(ivf)))

    def get_forcing_state(self):
        b = np.zeros(450)
        s, y = C.c_int(), C.c_int()
        iv = np.zeros(32, dtype=np.int32)
        self._chk(self.lib.orc_get_forcing_state(self.h, dp(b), C.byref(s), C.byref(y), ip(iv)))
        return b, s.value, y.value, iv

    def supfor_sth(self, v=None):
        """One supforSTH call (flowlib.f90:378-562) on the handle's forcing state.
        Returns the forcing table a1r..a3r as an array (3, 6, 5, 5): component, kx
        re/im entry, ky slot, kz slot (slots 1..5 hold k = 0, 1, 2, -2, -1).  When v
        (a spectral triple in the reference layout) is given, dt * a is added to it in
        place, as the reference adds it to the velocity."""
        a = np.zeros(450)
        if v is None:
            self._chk(self.lib.orc_supfor_sth(self.h, None, None, None, dp(a)))
        else:
            for x in v:
                assert x.flags.f_contiguous and x.dtype == np.float64
            self._chk(self.lib.orc_supfor_sth(self.h, *[dp(x) for x in v], dp(a)))
        return np.stack([a[150 * c:150 * (c + 1)].reshape((6, 5, 5), order="F") for c in range(3)])

    def set_droplets(self, yp, fvp=None, fdp=None, pertvel=None):
        yp = np.asfortranarray(yp, dtype=np.float64)
        cv = lambda x: None if x is None else np.asfortranarray(x, dtype=np.float64)
        self._chk(self.lib.orc_set_droplets(self.h, yp.shape[1], dp(yp), dp(cv(fvp)), dp(cv(fdp)), dp(cv(pertvel))))

    def get_droplets(self):
        cap = max(self.p.capacity if self.p.capacity > 0 else self.p.npart, 1)
        n = C.c_int()
        yp = np.zeros((8, cap), order="F"); fvp = np.zeros((3, 3, cap), order="F")
        fdp = np.zeros((3, 4, cap), order="F"); pv = np.zeros((3, cap), order="F")
        vp = np.zeros((3, cap), order="F"); dm = np.zeros((3, cap), order="F")
        self._chk(self.lib.orc_get_droplets(self.h, C.byref(n), dp(yp), dp(fvp), dp(fdp), dp(pv), dp(vp), dp(dm)))
        m = n.value
        return dict(yp=yp[:, :m], fvp=fvp[:, :, :m], fdp=fdp[:, :, :m], pertvel=pv[:, :m], vp=vp[:, :m], dmove=dm[:, :m])

    def set_new_particle(self, flag):
        self.lib.orc_set_new_particle(self.h, int(flag))

    def flow_only(self, on=True):
        self.lib.orc_flow_only(self.h, int(on))

    def step(self, n=1):
        return self._chk(self.lib.orc_step(self.h, n))

    def spectrum(self):
        ek = np.zeros(self.nek)
        self._chk(self.lib.orc_get_spectrum(self.h, dp(ek)))
        return ek

    def spectra(self):
        e, d, t = (np.zeros(self.nek) for _ in range(3))
        self._chk(self.lib.orc_get_spectra(self.h, dp(e), dp(d), dp(t)))
        return e, d, t

    def monitors(self, divstat=False):
        m = Monitors()
        self._chk(self.lib.orc_get_monitors(self.h, 1 if divstat else 0, C.byref(m)))
        return m

    def physical_velocity(self):
        n3 = self.N ** 3
        u = [np.zeros(n3) for _ in range(3)]
        self._chk(self.lib.orc_get_physical_velocity(self.h, *[dp(a) for a in u]))
        return [a.reshape(self.N, self.N, self.N) for a in u]

    def collision_counts(self):
        c = CollisionCounts()
        self._chk(self.lib.orc_get_collision_counts(self.h, C.byref(c)))
        return c

    def collision_bins(self):
        ns = self.p.nset
        nb = ns * ns * NBIN
        nir = np.zeros(nb, dtype=np.int32)
        b = [np.zeros(nb) for _ in range(7)]
        self._chk(self.lib.orc_get_collision_bins(self.h, ip(nir), *[dp(x) for x in b]))
        out = {"nirbin": nir.reshape((ns, ns, NBIN), order="F")}
        for k, a in zip(("wrbin", "wxbin", "wybin", "wzbin", "urbin", "thetabin", "wrtanbin"), b):
            out[k] = a.reshape((ns, ns, NBIN), order="F")
        return out

    def collision_pairs(self):
        n = C.c_int()
        self.lib.orc_get_collision_pairs(self.h, None, 0, C.byref(n))
        t = np.zeros((max(n.value, 1), 3), dtype=np.int64)
        self.lib.orc_get_collision_pairs(self.h, t.ctypes.data_as(C.POINTER(C.c_int64)), n.value, C.byref(n))
        return t[: n.value]

    def contact_records(self):
        """(m, 12) records of the pairs in contact, in visiting order (partlib.f90:1431-1437)."""
        n = C.c_int()
        self.lib.orc_get_contact_records(self.h, None, 0, C.byref(n))
        r = np.zeros((12, max(n.value, 1)), order="F")      # record-major in memory
        self.lib.orc_get_contact_records(self.h, dp(r), n.value, C.byref(n))
        return r.T[: n.value]

    def removal_list(self):
        n = C.c_int()
        self.lib.orc_get_removal_list(self.h, None, 0, C.byref(n))
        a = np.zeros(max(n.value, 1), dtype=np.int32)
        self.lib.orc_get_removal_list(self.h, ip(a), n.value, C.byref(n))
        return a[: n.value]

    def collision_detect(self, yp, dmove, fvp0, vnew, vp):
        a = [np.asfortranarray(x, dtype=np.float64) for x in (yp, dmove, fvp0, vnew, vp)]
        self._chk(self.lib.orc_collision_detect(self.h, a[0].shape[1], *[dp(x) for x in a]))

    def overlap_fix(self):
        return self._chk(self.lib.orc_overlap_fix(self.h))

    def init_droplets(self, seed=290710, radii=0, median_um=25.0, sigma_ln=0.4, rmin_um=5.0, rmax_um=60.0):
        return self._chk(self.lib.orc_init_droplets(self.h, seed, radii, median_um, sigma_ln, rmin_um, rmax_um))

    def scaling(self):
        out = np.zeros(16)
        self.lib.orc_get_scaling(self.h, dp(out))
        return dict(radius=out[:10].copy(), radmax=out[10], Sv_const=out[11], tau2=out[12],
                    d_coll=out[13], d_hdi=out[14], gravity_dns=out[15])

    def relocate_state(self):
        s, y, d = C.c_int(), C.c_int(), C.c_int()
        iv = np.zeros(32, dtype=np.int32)
        self.lib.orc_get_relocate_state(self.h, C.byref(s), C.byref(y), ip(iv), C.byref(d))
        return s.value, y.value, iv, d.value

    def hdi_solve(self, yp, vold, vp, pertvel):
        yp = np.asfortranarray(yp, dtype=np.float64)
        vold = np.asfortranarray(vold, dtype=np.float64)
        vp = np.asfortranarray(vp, dtype=np.float64)
        pv = np.array(pertvel, dtype=np.float64, order="F", copy=True)
        it, res = C.c_int(), C.c_double()
        st = self.lib.orc_hdi_solve(self.h, yp.shape[1], dp(yp), dp(vold), dp(vp), dp(pv), C.byref(it), C.byref(res))
        return st, pv, it.value, res.value


def fft_r2c(v):
    v = np.array(v, dtype=np.float64, order="F", copy=True)
    lib().orc_fft_r2c(v.shape[2], dp(v))
    return v


def fft_c2r(v):
    v = np.array(v, dtype=np.float64, order="F", copy=True)
    lib().orc_fft_c2r(v.shape[2], dp(v))
    return v
