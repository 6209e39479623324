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
