"""The flow path at 512^3, above the sizes the oracle can follow, held to properties
that do not depend on the size: the 3-D transform pair is an identity, the
FP32-mixed velocity stays within 1e-5 relative L2 of the FP64 one after 10 steps,
the field stays solenoidal, and the spectrum sums to the monitors' energy.  The
512-point plans and the z stage as separate kernels (the path from N = 512 up) run
nowhere else at their own size."""
import numpy as np
import pytest

from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

N = 512


def test_fft_roundtrip_512(ddns):
    # a band-limited field (no content on the Nyquist planes, which the transform
    # pair does not carry): products of cosines up to wavenumber 255
    rng = np.random.default_rng(7)
    L = N + 2
    x = 2 * np.pi * np.arange(N) / N
    phys = np.zeros((L, L, N), order="F")
    for k in ((0, 0, 0), (3, 100, 255), (1, 2, 3), (200, 17, 5), (255, 255, 1)):
        c = [np.cos(k[d] * x + rng.uniform(0, 2 * np.pi)) for d in range(3)]
        phys[:N, :N, :] += rng.standard_normal() * (c[0][:, None, None] * c[1][None, :, None] * c[2][None, None, :])
    for prec, tol in ((0, 1e-13), (1, 1e-5)):
        with ddns.Solver(ddns.default_params(N=N, precision=prec)) as s:
            back = s.fft_c2r(s.fft_r2c(phys))
        err = np.abs(back[:N, :N, :] - phys[:N, :N, :]).max() / np.abs(phys).max()
        assert err < tol, (prec, err)


def test_flow_512_fp32_mixed_against_fp64(ddns):
    """Deterministic forcing, k^4 exp(-2 (k/k0)^2) start (k0 = 4, u_rms = 1), 10 steps
    at a CFL number near 0.3."""
    kw = dict(N=N, forcing=1, dt=0.001, rnu=0.002)
    out = {}
    for prec in (0, 1):
        with ddns.Solver(ddns.default_params(precision=prec, **kw)) as s:
            s.init_flow(ddns.SPEC_K4_GAUSS, seed=123, k0=4.0, urms=1.0)
            s.step(10)
            v, _ = s.get_flow()
            m = s.monitors(divstat=True)
            ek = s.spectrum()
            out[prec] = (v, m, ek)
    v64, m64, ek64 = out[0]
    v32, m32, ek32 = out[1]
    assert rel_l2(v32, v64) < 1e-5
    for m, ek, tol in ((m64, ek64, 1e-10), (m32, ek32, 1e-5)):
        assert np.all(np.isfinite(ek)) and ek.sum() > 0
        assert abs(ek.sum() - m.qq) <= 1e-9 * m.qq          # total energy of monitorb.dat
        assert m.vdiv2 <= tol * m.vort2                     # solenoidal: rms div u against rms |w|


def test_split_z_stage_matches_fused_512(ddns, monkeypatch):
    """At 512^3 both z-stage paths fit (the separate kernels are the default there): they
    give the same field."""
    kw = dict(N=N, forcing=1, dt=0.001, rnu=0.002, precision=0)
    out = []
    for split in ("0", "1"):
        monkeypatch.setenv("DDNS_Z_SPLIT", split)
        with ddns.Solver(ddns.default_params(**kw)) as s:
            s.init_flow(ddns.SPEC_K4_GAUSS, seed=5, k0=4.0, urms=1.0)
            s.step(3)
            out.append(s.get_flow()[0])
    assert rel_l2(out[1], out[0]) < 1e-13


def test_split_z_stage_with_droplets_and_hi_512(ddns, monkeypatch):
    """A 512^3 step with droplets, HI and collisions beside the z stage, with the separate
    z kernels (the default at this size) and with the fused one: the same flow, and the
    droplets, their disturbance velocities and the collision counts agree."""
    kw = dict(N=N, forcing=1, dt=0.001, rnu=0.002, precision=0, npart=50000, radius=[40.0],
              hdi_included=1, hdi_trunc=20.0)
    out = []
    for split in ("1", "0"):
        monkeypatch.setenv("DDNS_Z_SPLIT", split)
        with ddns.Solver(ddns.default_params(**kw)) as s:
            s.init_flow(ddns.SPEC_K4_GAUSS, seed=11, k0=4.0, urms=1.0)
            s.init_droplets(seed=290710)
            s.step(3)
            c = s.collision_counts()
            out.append((s.get_flow()[0], s.get_droplets(), c.gmres_iters, list(c.icc)))
    (u1, d1, it1, icc1), (u0, d0, it0, icc0) = out
    print(f"512^3 with 50000 droplets: split against fused velocity {rel_l2(u1, u0):.2e}, GMRES {it1} / {it0}", flush=True)
    assert rel_l2(u1, u0) < 1e-13
    assert it1 == it0 >= 1 and icc1 == icc0
    assert np.abs(d1["yp"] - d0["yp"]).max() < 1e-10
    assert np.abs(d1["pertvel"] - d0["pertvel"]).max() < 1e-9 * max(1e-30, np.abs(d0["pertvel"]).max())


def test_fft_roundtrip_1024_fp64(ddns):
    """The largest FP64 plans: the plain 1024-point z transforms (whose padded tile and
    twiddles fill the 160 KiB of LDS exactly) and the y and x stages, through the
    transform pair on a field of three separable modes."""
    n = 1024
    L = n + 2
    x = 2 * np.pi * np.arange(n) / n
    phys = np.zeros((L, L, n), order="F")
    phys[:n, :n, :] = (np.cos(511 * x + 0.3)[:, None, None] + 0.5 * np.sin(3 * x + 1.1)[None, :, None]
                       + 0.25 * np.cos(400 * x)[None, None, :])
    with ddns.Solver(ddns.default_params(N=n, precision=0)) as s:
        spec = s.fft_r2c(phys)
        back = s.fft_c2r(spec)
    err = np.abs(back[:n, :n, :] - phys[:n, :n, :]).max()
    print(f"1024^3 FP64 transform pair: max error {err:.2e}", flush=True)
    assert err < 1e-12
    # the forward transform holds the three modes and nothing else: at most two complex
    # coefficients (four real entries) per mode
    nz = int(np.count_nonzero(np.abs(spec) > 1e-9))
    assert 3 <= nz <= 12, nz


def test_flow_1024_fp32_mixed():
    """The largest grid (1024^3, the 5M-droplet configuration's flow) through the device
    initial field and two steps: the z stage as separate kernels, 1024-point plans.  The
    start holds 3/2 u_rms^2 exactly; after a step the deterministic forcing has set the
    energies of shells 1 and 2 to force(1:2) (supforDET, flowlib.f90:195-229), and the
    field stays solenoidal."""
    import mpi_2d_domain_decomp_turb_particles_amd as dd
    p = dd.default_params(N=1024, precision=dd.FP32_MIXED, dt=0.00025, rnu=0.0008, forcing=1)
    with dd.Solver(p) as s:
        s.init_flow(dd.SPEC_K4_GAUSS, seed=123, k0=4.0, urms=1.0)
        m0 = s.monitors(divstat=True)
        assert abs(m0.qq - 1.5) < 1e-5
        s.step(2)
        m = s.monitors(divstat=True)
        ek = s.spectrum()
    assert np.all(np.isfinite(ek)) and abs(ek.sum() - m.qq) <= 1e-6 * m.qq
    assert abs(ek[0] - p.force[0]) <= 1e-5 * p.force[0] and abs(ek[1] - p.force[1]) <= 1e-5 * p.force[1]
    assert m.vdiv2 <= 1e-5 * m.vort2
