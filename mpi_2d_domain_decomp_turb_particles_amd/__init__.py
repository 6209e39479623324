"""MI355X-native DNS + droplet time step (HIP kernels behind a C ABI).

The reference's per-timestep hot path -- the pseudo-spectral Navier-Stokes step of
turb.f90 / mpifft.f90 / flowlib.f90 and the droplet step of partlib.f90 -- runs
on one GPU in ``libddns.so``.  This module is the Python host mirror of that C ABI
(include/ddns.h); the Fortran host driver in ``fortran/`` is the other one.

Arrays cross the boundary in the reference's own layouts:
  * spectral fields v(N+2, N+2, N), Fortran order: kz, kx re/im pairs, ky;
  * droplets yp(8, n), fvp(3, 0:2, n), fdp(3, 0:3, n), pertvel(3, n).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

from . import _abi
from ._abi import (  # noqa: F401  (re-exported)
    DDNS_ERR_ARG, DDNS_ERR_CAPACITY, DDNS_ERR_DOMAIN, DDNS_ERR_DROPLET_LOSS, DDNS_ERR_HIP,
    DDNS_ERR_IO, DDNS_ERR_NEIGHBOUR_OVERFLOW, DDNS_ERR_NO_DEVICE, DDNS_ERR_OVERLAP, DDNS_OK,
    DDNS_WARN_GMRES, FP32_MIXED, FP64, MAXSET, NBIN, RADII_LOGNORMAL, RADII_SETS,
    SPEC_K4_GAUSS, SPEC_KOLMOGOROV, CollisionCounts, Params, Timing, default_params,
)

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libddns.so"

_lib = None


class DDNSError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ddns error {code}: {msg}")
        self.code = code


def load_library(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load libddns.so (built in-tree by build.py).  There is no fallback: without
    the HIP library nothing in this package can run."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise ImportError(f"{p} is missing: build it with `python -m mpi_2d_domain_decomp_turb_particles_amd.build`")
    lib = C.CDLL(str(p))
    for name, (res, args) in _abi.SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _dp(a):
    if a is None:
        return None
    if a.dtype != np.float64 or not a.flags["F_CONTIGUOUS"]:
        # the C ABI takes the reference's column-major arrays
        raise TypeError("expected a Fortran-contiguous float64 array")
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_int))


def spectral_zeros(N: int) -> np.ndarray:
    """A spectral array in the reference layout v(N+2, N+2, N) (Fortran order)."""
    return np.zeros((N + 2, N + 2, N), dtype=np.float64, order="F")


def read_parameter_file(path: str | os.PathLike) -> Params:
    """Positional parameter.d reader (turb.f90:101-166)."""
    lib = load_library()
    p = Params()
    st = lib.ddns_read_parameter_file(str(path).encode(), C.byref(p))
    if st != 0:
        raise DDNSError(st, lib.ddns_last_error(None).decode())
    return p


class Solver:
    """One periodic domain resident on the current GPU."""

    def __init__(self, params: Params):
        self.lib = load_library()
        self.p = params
        self.N = params.nx
        h = C.c_void_p()
        st = self.lib.ddns_create(C.byref(params), C.byref(h))
        if st != 0:
            raise DDNSError(st, self.lib.ddns_last_error(None).decode())
        self.h = h
        self.nek = self.lib.ddns_nek(h)

    # -- lifetime
    def close(self):
        if self.h:
            self.lib.ddns_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st: int) -> int:
        if st < 0:
            raise DDNSError(st, self.lib.ddns_last_error(self.h).decode())
        return st

    # -- flow
    def set_flow(self, v, o=None):
        v = [np.asfortranarray(x, dtype=np.float64) for x in v]
        o = [np.asfortranarray(x, dtype=np.float64) for x in o] if o is not None else [None] * 3
        self._check(self.lib.ddns_set_flow(self.h, *[_dp(x) for x in v], *[_dp(x) for x in o]))

    def get_flow(self):
        arrs = [spectral_zeros(self.N) for _ in range(6)]
        self._check(self.lib.ddns_get_flow(self.h, *[_dp(a) for a in arrs]))
        return arrs[:3], arrs[3:]

    def init_flow(self, kind: int = SPEC_KOLMOGOROV, seed: int = 123, k0: float = 4.0, urms: float = 1.0):
        self._check(self.lib.ddns_init_flow(self.h, kind, seed, k0, urms))

    def set_forcing_state(self, b, iseedf: int, iyf: int, ivf):
        b = np.ascontiguousarray(b, dtype=np.float64)
        ivf = np.ascontiguousarray(ivf, dtype=np.int32)
        self._check(self.lib.ddns_set_forcing_state(self.h, _dp(b), iseedf, iyf, _ip(ivf)))

    def get_forcing_state(self):
        b = np.zeros(450)
        s, y = C.c_int(), C.c_int()
        iv = np.zeros(32, dtype=np.int32)
        self._check(self.lib.ddns_get_forcing_state(self.h, _dp(b), C.byref(s), C.byref(y), _ip(iv)))
        return b, s.value, y.value, iv

    def physical_velocity(self):
        n3 = self.N ** 3
        u = [np.zeros(n3) for _ in range(3)]
        self._check(self.lib.ddns_get_physical_velocity(self.h, *[_dp(a) for a in u]))
        return [a.reshape(self.N, self.N, self.N) for a in u]   # [z][y][x]

    def spectrum(self):
        ek = np.zeros(self.nek)
        self._check(self.lib.ddns_get_spectrum(self.h, _dp(ek)))
        return ek

    def spectra(self):
        """Energy, dissipation and transfer spectra (spectrum.dat / transfer.dat)."""
        e, d, t = (np.zeros(self.nek) for _ in range(3))
        self._check(self.lib.ddns_get_spectra(self.h, _dp(e), _dp(d), _dp(t)))
        return e, d, t

    def monitors(self, divstat: bool = False) -> "_abi.Monitors":
        """POSTPROC statistics (monitora.dat, monitorb.dat, monitorc.dat)."""
        m = _abi.Monitors()
        self._check(self.lib.ddns_get_monitors(self.h, 1 if divstat else 0, C.byref(m)))
        return m

    # -- droplets
    def set_droplets(self, yp, fvp=None, fdp=None, pertvel=None):
        yp = np.asfortranarray(yp, dtype=np.float64)
        n = yp.shape[1]
        fvp = None if fvp is None else np.asfortranarray(fvp, dtype=np.float64)
        fdp = None if fdp is None else np.asfortranarray(fdp, dtype=np.float64)
        pertvel = None if pertvel is None else np.asfortranarray(pertvel, dtype=np.float64)
        self._check(self.lib.ddns_set_droplets(self.h, n, _dp(yp), _dp(fvp), _dp(fdp), _dp(pertvel)))

    def get_droplets(self):
        cap = max(self.p.capacity if self.p.capacity > 0 else self.p.npart, 1)
        n = C.c_int()
        yp = np.zeros((8, cap), order="F")
        fvp = np.zeros((3, 3, cap), order="F")
        fdp = np.zeros((3, 4, cap), order="F")
        pv = np.zeros((3, cap), order="F")
        vp = np.zeros((3, cap), order="F")
        dm = np.zeros((3, cap), order="F")
        self._check(self.lib.ddns_get_droplets(self.h, C.byref(n), _dp(yp), _dp(fvp), _dp(fdp), _dp(pv), _dp(vp), _dp(dm)))
        m = n.value
        return dict(yp=yp[:, :m], fvp=fvp[:, :, :m], fdp=fdp[:, :, :m], pertvel=pv[:, :m], vp=vp[:, :m], dmove=dm[:, :m])

    def init_droplets(self, seed=290710, radii=RADII_SETS, median_um=25.0, sigma_ln=0.4, rmin_um=5.0, rmax_um=60.0):
        self._check(self.lib.ddns_init_droplets(self.h, seed, radii, median_um, sigma_ln, rmin_um, rmax_um))

    def overlap_fix(self):
        self._check(self.lib.ddns_overlap_fix(self.h))

    # -- step
    def step(self, nsteps: int = 1) -> int:
        return self._check(self.lib.ddns_step(self.h, nsteps))

    def collision_counts(self) -> CollisionCounts:
        c = CollisionCounts()
        self._check(self.lib.ddns_get_collision_counts(self.h, C.byref(c)))
        return c

    def collision_bins(self):
        ns = self.p.nset
        nb = ns * ns * NBIN
        nir = np.zeros(nb, dtype=np.int32)
        b = [np.zeros(nb) for _ in range(7)]
        self._check(self.lib.ddns_get_collision_bins(self.h, _ip(nir), *[_dp(x) for x in b]))
        shape = (ns, ns, NBIN)
        out = {"nirbin": nir.reshape(shape, order="F")}
        for k, a in zip(("wrbin", "wxbin", "wybin", "wzbin", "urbin", "thetabin", "wrtanbin"), b):
            out[k] = a.reshape(shape, order="F")
        return out

    def collision_pairs(self):
        n = C.c_int()
        self._check(self.lib.ddns_get_collision_pairs(self.h, None, 0, C.byref(n)))
        if n.value == 0:
            return np.zeros((0, 3), dtype=np.int64)
        t = np.zeros((n.value, 3), dtype=np.int64)
        self._check(self.lib.ddns_get_collision_pairs(self.h, t.ctypes.data_as(C.POINTER(C.c_int64)), n.value, C.byref(n)))
        return t

    def contact_records(self):
        """(m, 12) records of the last step's pairs in contact, buffer(1:12, ij) of
        COLLISION_DET (partlib.f90:1431-1437), in the one-rank visiting order."""
        n = C.c_int()
        self._check(self.lib.ddns_get_contact_records(self.h, None, 0, C.byref(n)))
        r = np.zeros((12, n.value), order="F")        # record-major in memory
        if n.value:
            self._check(self.lib.ddns_get_contact_records(self.h, _dp(r), n.value, C.byref(n)))
        return r.T

    def tuning(self):
        """Tile shapes chosen at set-up: dict with the y-stage tile width and its default,
        the fused z stage's threads and their default, and the timings of the candidates
        (ms; 0 where a candidate was not timed)."""
        sh = np.zeros(4, dtype=np.int32)
        ms = np.zeros(6)
        self._check(self.lib.ddns_get_tuning(self.h, sh.ctypes.data_as(C.POINTER(C.c_int)), _dp(ms)))
        return {"ky_tc": int(sh[0]), "ky_tc_default": int(sh[1]), "kz_nt": int(sh[2]),
                "kz_nt_default": int(sh[3]), "ky_ms": ms[:3].copy(), "kz_ms": ms[3:].copy()}

    def timing(self) -> Timing:
        t = Timing()
        self._check(self.lib.ddns_get_timing(self.h, C.byref(t)))
        return t

    # -- kernel level
    def fft_r2c(self, v):
        v = np.array(v, dtype=np.float64, order="F", copy=True)
        self._check(self.lib.ddns_fft_r2c(self.h, _dp(v)))
        return v

    def fft_c2r(self, v):
        v = np.array(v, dtype=np.float64, order="F", copy=True)
        self._check(self.lib.ddns_fft_c2r(self.h, _dp(v)))
        return v

    def collision_detect(self, yp, dmove, fvp0, vnew, vp):
        a = [np.asfortranarray(x, dtype=np.float64) for x in (yp, dmove, fvp0, vnew, vp)]
        self._check(self.lib.ddns_collision_detect(self.h, a[0].shape[1], *[_dp(x) for x in a]))

    def profile_flow(self, k: int = 5):
        ms = np.zeros(7)
        self._check(self.lib.ddns_profile_flow(self.h, k, _dp(ms), 7))
        return dict(zip(("y_inv", "x_fused", "y_fwd", "pre", "z_stage", "sym", "step"), ms))

    def profile_copy(self, reps: int = 20):
        """Mean time (ms) and bytes moved per launch of a plain HBM copy kernel."""
        ms, by = C.c_double(), C.c_double()
        self._check(self.lib.ddns_profile_copy(self.h, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value


def stokes_radius_um(p: Params, st: float) -> float:
    """Droplet radius (um) whose Stokes number tau_p / tau_k is st, with the response time
    and Kolmogorov time of PART_INIT's scaling (partlib.f90:104-150):
    tau_p = 2 rho_w a^2 / (9 rho_a nu), tau_k = sqrt(nu / eps), a in cm."""
    tauk = np.sqrt(p.viscosity_CLOUD / p.ediss_CLOUD)
    a_cm = np.sqrt(st * tauk * 9.0 * p.rho_air * p.viscosity_CLOUD / (2.0 * p.rho_water))
    return float(a_cm * 1.0e4)


def batched_realisations(R: int = 48, N: int = 64, precision: int = FP32_MIXED, seed: int = 2024,
                         nmin: int = 5000, nmax: int = 200000, st_min: float = 0.2, st_max: float = 3.0,
                         mean_neighbours: float = 24.0, hdi: bool = True):
    """Parameter sets of R independent realisations for a Batch: the reference's N^3
    defaults, seeds that differ per realisation (flow iseed_m0, forcing iseedf, relocation
    draws), deterministic forcing on the even realisations and stochastic (OU) forcing on
    the odd ones, droplet counts drawn uniformly from [nmin, nmax], one droplet size per
    realisation with Stokes numbers evenly spaced over [st_min, st_max].  With HI, each
    realisation's truncation radius is 50 droplet radii, or less where that would give a
    droplet more than mean_neighbours neighbours on average, so every HI row stays well
    inside the reference's limit of 100 neighbours."""
    rng = np.random.default_rng(seed)
    counts = rng.integers(nmin, nmax + 1, size=R)
    sts = np.linspace(st_min, st_max, R) if R > 1 else np.array([st_min])
    vol = (2.0 * np.pi) ** 3
    out = []
    for r in range(R):
        p = default_params(N=N, precision=precision)
        p.forcing = 1 if r % 2 == 0 else 2
        p.iseed_m0 = 1000 + r
        p.iseedf = 232300 + 17 * r
        p.relocate_seed = r
        p.npart = int(counts[r])
        p.nset = 1
        p.radius[0] = stokes_radius_um(p, float(sts[r]))
        if hdi:
            etk_cloud = (p.viscosity_CLOUD ** 3 / p.ediss_CLOUD) ** 0.25
            a_code = p.radius[0] * (p.etk_DNS / etk_cloud) / 1.0e4
            reach = (mean_neighbours * vol / (p.npart * 4.0 / 3.0 * np.pi)) ** (1.0 / 3.0)
            p.hdi_included, p.nonoverlap = 1, 1
            p.hdi_trunc = float(min(50.0, reach / a_code))
        out.append(p)
    return out


class _Member(Solver):
    """One realisation of a Batch: a Solver over the member handle, owned by the batch
    (closing it does nothing; stepping it is refused by the library)."""

    def __init__(self, batch: "Batch", r: int):
        self.lib = batch.lib
        self.p = batch.params[r]
        self.N = self.p.nx
        h = self.lib.ddns_batch_member(batch.b, r)
        if not h:
            raise DDNSError(DDNS_ERR_ARG, f"no realisation {r}")
        self.h = C.c_void_p(h)
        self.nek = self.lib.ddns_nek(self.h)
        self._batch = batch

    def close(self):
        self.h = None


class Batch:
    """R independent realisations of one grid, precision, dt and viscosity advanced
    together: every flow stage is one launch over all of them, and each realisation's
    droplet block runs as a Solver's does.  ``member(r)`` gives realisation r's state
    (flow, forcing, droplets, collision statistics) as a Solver."""

    def __init__(self, params):
        self.lib = load_library()
        self.params = list(params)
        arr = (Params * len(self.params))(*self.params)
        b = C.c_void_p()
        st = self.lib.ddns_batch_create(arr, len(self.params), C.byref(b))
        if st != 0:
            raise DDNSError(st, self.lib.ddns_batch_last_error(None).decode())
        self.b = b
        self.R = len(self.params)
        self._members = [_Member(self, r) for r in range(self.R)]

    def member(self, r: int) -> Solver:
        return self._members[r]

    def __len__(self):
        return self.R

    def close(self):
        if self.b:
            for m in self._members:
                m.close()
            self.lib.ddns_batch_destroy(self.b)
            self.b = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st: int) -> int:
        if st < 0:
            raise DDNSError(st, self.lib.ddns_batch_last_error(self.b).decode())
        return st

    def step(self, nsteps: int = 1) -> int:
        return self._check(self.lib.ddns_batch_step(self.b, nsteps))

    def timing(self) -> Timing:
        t = Timing()
        self._check(self.lib.ddns_batch_get_timing(self.b, C.byref(t)))
        return t

    def tuning(self):
        sh = np.zeros(4, dtype=np.int32)
        ms = np.zeros(6)
        self._check(self.lib.ddns_batch_get_tuning(self.b, sh.ctypes.data_as(C.POINTER(C.c_int)), _dp(ms)))
        return {"ky_tc": int(sh[0]), "ky_tc_default": int(sh[1]), "kz_nt": int(sh[2]),
                "kz_nt_default": int(sh[3]), "ky_ms": ms[:3].copy(), "kz_ms": ms[3:].copy()}
