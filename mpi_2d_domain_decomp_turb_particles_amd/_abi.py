"""ctypes mirror of include/ddns.h (structs, constants, function signatures)."""
from __future__ import annotations

import ctypes as C

DDNS_OK = 0
DDNS_ERR_ARG = -1
DDNS_ERR_HIP = -2
DDNS_ERR_OVERLAP = -3
DDNS_ERR_NEIGHBOUR_OVERFLOW = -4
DDNS_ERR_DROPLET_LOSS = -5
DDNS_ERR_DOMAIN = -6
DDNS_ERR_CAPACITY = -7
DDNS_ERR_IO = -8
DDNS_ERR_NO_DEVICE = -9
DDNS_WARN_GMRES = 1

FP64 = 0
FP32_MIXED = 1
NBIN = 180
MAXSET = 10
MAXNEI = 100
SPEC_KOLMOGOROV = 0
SPEC_K4_GAUSS = 1
RADII_SETS = 0
RADII_LOGNORMAL = 1


class Params(C.Structure):
    """ddns_params: parameter.d fields in read order (turb.f90:103-164), then the
    values the reference hard-codes."""

    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("nprocY", C.c_int), ("nstep", C.c_int),
        ("dt", C.c_double), ("rnu", C.c_double), ("time", C.c_double),
        ("new_flow", C.c_int), ("idp", C.c_int), ("iseed_m0", C.c_int), ("forcing", C.c_int),
        ("force", C.c_double * 2),
        ("iseedf", C.c_int), ("ieout", C.c_int), ("ispec", C.c_int), ("divstat", C.c_int),
        ("rho_water", C.c_double), ("rho_air", C.c_double),
        ("etk_DNS", C.c_double), ("tauk_DNS", C.c_double),
        ("st", C.c_double), ("ediss_CLOUD", C.c_double), ("viscosity_CLOUD", C.c_double),
        ("gravity", C.c_double),
        ("new_particle", C.c_int), ("nonoverlap", C.c_int), ("hdi_included", C.c_int),
        ("hdi_trunc", C.c_double),
        ("npart", C.c_int), ("ndrag", C.c_int), ("nset", C.c_int), ("pad0", C.c_int),
        ("radius", C.c_double * MAXSET),
        ("precision", C.c_int), ("nstep_grow", C.c_int),
        ("ou_tf", C.c_double), ("ou_var", C.c_double),
        ("gmres_maxiter", C.c_int), ("capacity", C.c_int),
        ("gmres_tol", C.c_double),
        ("relocate_seed", C.c_int), ("istep0", C.c_int),
    ]


class CollisionCounts(C.Structure):
    _fields_ = [
        ("icc", C.c_int * 3), ("ncoll", C.c_int), ("npairs", C.c_int), ("nset", C.c_int),
        ("colltype", C.c_int * (MAXSET * MAXSET)), ("nir", C.c_int * (MAXSET * MAXSET)),
        ("pairs_tested", C.c_int64), ("pairs_in_dstat", C.c_int64),
        ("noverlap_sweeps", C.c_int), ("noverlap_relocated", C.c_int),
        ("gmres_iters", C.c_int), ("pad1", C.c_int), ("gmres_residual", C.c_double),
    ]


class Monitors(C.Structure):
    """ddns_monitors: POSTPROC statistics (postprocessing.f90:76-272)."""
    _fields_ = [(f, C.c_double) for f in (
        "eta", "tmse", "xintls", "tk", "et", "xl", "uprime",
        "qq", "dissp", "ree", "resol", "cfl", "vskew", "vflat", "pmax",
        "vdiv1", "vdiv2", "vort1", "vort2", "divvt")]


class Timing(C.Structure):
    _fields_ = [
        ("flow_ms", C.c_double), ("interp_ms", C.c_double), ("cells_ms", C.c_double),
        ("hdi_ms", C.c_double), ("advance_ms", C.c_double), ("colldet_ms", C.c_double),
        ("overlap_ms", C.c_double), ("total_ms", C.c_double), ("nsteps", C.c_int), ("pad", C.c_int),
        ("flow_kernel_ms", C.c_double * 6),
    ]


def default_params(**over) -> Params:
    """Reference defaults (the 64^3 parameter.d and the hard-coded constants);
    keyword arguments override fields (``force``/``radius`` take sequences)."""
    p = Params()
    p.nx = p.ny = p.nz = 64
    p.nprocY = 1
    p.nstep = 10
    p.dt = 0.001
    p.rnu = 0.026
    p.time = 0.0
    p.new_flow = 1
    p.idp = 0
    p.iseed_m0 = 123
    p.forcing = 1
    p.force[0], p.force[1] = 1.242477, 0.391356
    p.iseedf = 232300
    p.ieout = 10
    p.ispec = 1000
    p.divstat = 1
    p.rho_water = 1.0
    p.rho_air = 0.001
    p.etk_DNS, p.tauk_DNS = 0.0755, 0.2195
    p.st = 2.0
    p.ediss_CLOUD = 400.0
    p.viscosity_CLOUD = 0.17
    p.gravity = 980.0
    p.new_particle = 1
    p.nonoverlap = 0
    p.hdi_included = 0
    p.hdi_trunc = 50.0
    p.npart = 0
    p.ndrag = 0
    p.nset = 1
    p.radius[0] = 20.0
    p.precision = FP64
    p.nstep_grow = 40000
    p.ou_tf = 0.038
    p.ou_var = 447.31
    p.gmres_maxiter = 50
    p.capacity = 0
    p.gmres_tol = 1.0e-6
    p.relocate_seed = 0
    p.istep0 = 0
    for k, v in over.items():
        if k == "N":
            p.nx = p.ny = p.nz = int(v)
        elif k == "force":
            p.force[0], p.force[1] = float(v[0]), float(v[1])
        elif k == "radius":
            for i, r in enumerate(v):
                p.radius[i] = float(r)
            p.nset = len(v)
        else:
            if not hasattr(p, k):
                raise AttributeError(f"ddns_params has no field {k}")
            setattr(p, k, v)
    if p.hdi_included:
        p.nonoverlap = 1
    return p


_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_I64 = C.POINTER(C.c_int64)

# name -> (restype, argtypes); every entry point of include/ddns.h
SIGNATURES = {
    "ddns_params_default": (C.c_int, [C.POINTER(Params)]),
    "ddns_read_parameter_file": (C.c_int, [C.c_char_p, C.POINTER(Params)]),
    "ddns_create": (C.c_int, [C.POINTER(Params), C.POINTER(_P)]),
    "ddns_destroy": (C.c_int, [_P]),
    "ddns_last_error": (C.c_char_p, [_P]),
    "ddns_nek": (C.c_int, [_P]),
    "ddns_set_flow": (C.c_int, [_P, _D, _D, _D, _D, _D, _D]),
    "ddns_get_flow": (C.c_int, [_P, _D, _D, _D, _D, _D, _D]),
    "ddns_set_forcing_state": (C.c_int, [_P, _D, C.c_int, C.c_int, _I]),
    "ddns_get_forcing_state": (C.c_int, [_P, _D, _I, _I, _I]),
    "ddns_init_flow": (C.c_int, [_P, C.c_int, C.c_int, C.c_double, C.c_double]),
    "ddns_get_physical_velocity": (C.c_int, [_P, _D, _D, _D]),
    "ddns_get_spectrum": (C.c_int, [_P, _D]),
    "ddns_get_spectrum_device": (C.c_int, [_P, _P, _P]),
    "ddns_get_spectra": (C.c_int, [_P, _D, _D, _D]),
    "ddns_get_monitors": (C.c_int, [_P, C.c_int, C.POINTER(Monitors)]),
    "ddns_set_droplets": (C.c_int, [_P, C.c_int, _D, _D, _D, _D]),
    "ddns_get_droplets": (C.c_int, [_P, _I, _D, _D, _D, _D, _D, _D]),
    "ddns_init_droplets": (C.c_int, [_P, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]),
    "ddns_overlap_fix": (C.c_int, [_P]),
    "ddns_step": (C.c_int, [_P, C.c_int]),
    "ddns_get_collision_counts": (C.c_int, [_P, C.POINTER(CollisionCounts)]),
    "ddns_get_collision_bins": (C.c_int, [_P, _I, _D, _D, _D, _D, _D, _D, _D]),
    "ddns_get_contact_records": (C.c_int, [_P, _D, C.c_int, _I]),
    "ddns_get_collision_pairs": (C.c_int, [_P, _I64, C.c_int, _I]),
    "ddns_get_timing": (C.c_int, [_P, C.POINTER(Timing)]),
    "ddns_get_tuning": (C.c_int, [_P, _I, _D]),
    "ddns_fft_r2c": (C.c_int, [_P, _D]),
    "ddns_fft_c2r": (C.c_int, [_P, _D]),
    "ddns_collision_detect": (C.c_int, [_P, C.c_int, _D, _D, _D, _D, _D]),
    "ddns_profile_flow": (C.c_int, [_P, C.c_int, _D, C.c_int]),
    "ddns_profile_copy": (C.c_int, [_P, C.c_int, _D, _D]),
    "ddns_batch_create": (C.c_int, [C.POINTER(Params), C.c_int, C.POINTER(_P)]),
    "ddns_batch_destroy": (C.c_int, [_P]),
    "ddns_batch_size": (C.c_int, [_P]),
    "ddns_batch_member": (_P, [_P, C.c_int]),
    "ddns_batch_step": (C.c_int, [_P, C.c_int]),
    "ddns_batch_get_timing": (C.c_int, [_P, C.POINTER(Timing)]),
    "ddns_batch_get_tuning": (C.c_int, [_P, _I, _D]),
    "ddns_batch_last_error": (C.c_char_p, [_P]),
}
