"""GPU parity against the FP64 oracle, through the C ABI (libddns.so).

Tolerances: FP64 mode uses different summation orders than the oracle (FFT
butterflies, reductions), so fields agree to ~1e-12 relative; FP32-mixed mode is
held to the project's budget of 1e-5 relative L2 after 10 steps.  Integer outputs
(collision pair lists, counts) must be identical.
"""
import numpy as np
import pytest

from tests.helpers import random_droplets, random_spectral, rel_l2

pytestmark = pytest.mark.gpu


def _params(ddns, **kw):
    return ddns.default_params(**kw)


@pytest.mark.parametrize("N", [32, 64])
@pytest.mark.parametrize("prec", [0, 1])
def test_fft_roundtrip_matches_oracle(ddns, oracle_mod, N, prec):
    O = oracle_mod
    rng = np.random.default_rng(N + prec)
    L = N + 2
    phys = np.zeros((L, L, N), order="F")
    phys[:N, :N, :] = rng.standard_normal((N, N, N))
    with ddns.Solver(_params(ddns, N=N, precision=prec)) as s:
        g = s.fft_r2c(phys)
        o = O.fft_r2c(phys)
        keep = np.zeros_like(g, dtype=bool)
        keep[:N, :N, :] = True                  # the GPU does not store kx = N/2
        err = np.abs(g[keep] - o[keep]).max() / np.abs(o[keep]).max()
        assert err < (1e-13 if prec == 0 else 2e-6), err
        o2 = o.copy()
        o2[:, N:, :] = 0
        gb = s.fft_c2r(o2)
        ob = O.fft_c2r(o2)
        err = np.abs(gb - ob).max() / np.abs(ob).max()
        assert err < (1e-13 if prec == 0 else 2e-6), err


@pytest.mark.parametrize("forcing", [0, 1, 2])
@pytest.mark.parametrize("prec", [0, 1])
def test_flow_ten_steps_match_oracle(ddns, oracle_mod, forcing, prec):
    N = 32
    p = _params(ddns, N=N, forcing=forcing, precision=prec, dt=0.002, rnu=0.02)
    v = random_spectral(N, seed=3 + forcing, kmax=6)
    with ddns.Solver(p) as s, oracle_mod.Oracle(_params(ddns, N=N, forcing=forcing, dt=0.002, rnu=0.02)) as o:
        s.set_flow(v, v)
        o.set_flow(v, v)
        o.flow_only(True)
        s.step(10)
        o.step(10)
        gv, go = s.get_flow()
        ov, oo = o.get_flow()
        e = rel_l2(gv, ov)
        tol = 1e-11 if prec == 0 else 1e-5
        assert e < tol, f"velocity rel L2 {e}"
        eo = rel_l2(go, oo)
        assert eo < (1e-10 if prec == 0 else 1e-4), f"history rel L2 {eo}"
        ek_g, ek_o = s.spectrum(), o.spectrum()
        se = np.abs(ek_g - ek_o).max() / np.abs(ek_o).max()
        assert se < (1e-11 if prec == 0 else 1e-5), se
        if forcing == 2:
            bg, sg, yg, ivg = s.get_forcing_state()
            bo, so, yo, ivo = o.get_forcing_state()
            assert sg == so and yg == yo and np.array_equal(ivg, ivo)
            assert np.abs(bg - bo).max() < 1e-12 * max(1.0, np.abs(bo).max())


@pytest.mark.parametrize("kind", [1, 0])          # SPEC_K4_GAUSS, SPEC_KOLMOGOROV
@pytest.mark.parametrize("prec", [0, 1])
def test_initial_field_matches_oracle(ddns, oracle_mod, kind, prec):
    """initial_flow on the device against the oracle's restatement (same counter-based
    draws; the transcendental functions and the shell sums round differently)."""
    N = 64
    with ddns.Solver(_params(ddns, N=N, precision=prec)) as s, oracle_mod.Oracle(_params(ddns, N=N)) as o:
        s.init_flow(kind, seed=77, k0=4.0, urms=1.0)
        o.init_flow(kind, 77, 4.0, 1.0)
        gv, go = s.get_flow()
        ov, oo = o.get_flow()
        tol = 1e-12 if prec == 0 else 1e-6
        assert rel_l2(gv, ov) < tol and rel_l2(go, oo) < tol
        s.init_flow(-1, seed=77, k0=4.0, urms=1.0)         # at rest
        gv, go = s.get_flow()
        assert all(np.all(x == 0) for x in gv + go)


@pytest.mark.parametrize("prec", [0, 1])
def test_flow_statistics_match_oracle(ddns, oracle_mod, prec):
    """POSTPROC and TRANSPOSTPROC statistics (postprocessing.f90:76-315) of one u-hat:
    the energy, dissipation and transfer spectra and every monitora/b/c field.  FP64:
    summation orders differ, 1e-10 relative; FP32-mixed: the physical fields are
    float, 1e-4 relative (skewness 2e-3 absolute, the divergence of a solenoidal
    field only to its float rounding)."""
    N = 32
    v = random_spectral(N, seed=3, kmax=7)          # a field whose triads carry transfer
    kw = dict(N=N, forcing=0, dt=0.003, rnu=0.02)
    with ddns.Solver(_params(ddns, precision=prec, **kw)) as s, oracle_mod.Oracle(_params(ddns, **kw)) as o:
        s.set_flow(v, v)
        o.set_flow(v, v)
        gs, os_ = s.spectra(), o.spectra()
        gm, om = s.monitors(divstat=True), o.monitors(divstat=True)
    rt = 1e-10 if prec == 0 else 1e-4
    assert np.abs(os_[2]).max() > 0.1
    for g, r in zip(gs, os_):
        assert np.abs(g - r).max() <= rt * np.abs(r).max(), (g, r)
    assert abs(os_[2].sum()) < 1e-10 * np.abs(os_[2]).sum()
    for f in ("eta", "tmse", "xintls", "tk", "et", "xl", "uprime", "qq", "dissp", "ree", "resol", "cfl",
              "vflat", "pmax", "vort1", "vort2"):
        a, b = getattr(gm, f), getattr(om, f)
        assert abs(a - b) <= rt * abs(b), (f, a, b)
    assert abs(gm.vskew - om.vskew) <= (1e-10 if prec == 0 else 2e-3), (gm.vskew, om.vskew)
    div_tol = 1e-9 if prec == 0 else 1e-3
    assert abs(gm.vdiv1) < div_tol and gm.vdiv2 < div_tol and abs(gm.divvt) < div_tol


def test_physical_velocity_matches_oracle(ddns, oracle_mod):
    N = 32
    v = random_spectral(N, seed=11, kmax=6)
    p = _params(ddns, N=N, forcing=1)
    with ddns.Solver(p) as s, oracle_mod.Oracle(p) as o:
        s.set_flow(v, v)
        o.set_flow(v, v)
        o.flow_only(True)
        s.step(1)
        o.step(1)
        ug = s.physical_velocity()
        uo = o.physical_velocity()
        assert rel_l2(ug, uo) < 1e-12


def _collision_inputs(n, seed, radius, dt=0.002):
    rng = np.random.default_rng(seed)
    yp = random_droplets(n, seed, radius)
    v0 = rng.normal(0.0, 1.0, size=(3, n))
    vn = v0 + rng.normal(0.0, 0.5, size=(3, n))
    dmove = dt / 24.0 * (9 * vn + 19 * v0 - 5 * v0 + v0)
    yp[3:6, :] = vn
    vp = rng.normal(0.0, 1.0, size=(3, n))
    return yp, np.asfortranarray(dmove), np.asfortranarray(v0), np.asfortranarray(vn), np.asfortranarray(vp)


@pytest.mark.parametrize("hdi", [0, 1])
def test_collision_pairs_bit_identical(ddns, oracle_mod, hdi):
    N = 64
    p = _params(ddns, N=N, npart=30000, nset=2, radius=[60.0, 90.0], dt=0.01, hdi_included=hdi, hdi_trunc=8.0)
    with oracle_mod.Oracle(p) as o, ddns.Solver(p) as s:
        rad = o.scaling()["radius"][:2]
        yp, dm, v0, vn, vp = _collision_inputs(30000, 7, rad, dt=0.01)
        o.collision_detect(yp, dm, v0, vn, vp)
        s.collision_detect(yp, dm, v0, vn, vp)
        po, pg = o.collision_pairs(), s.collision_pairs()
        assert len(po) > 20
        assert np.array_equal(po, pg)
        co, cg = o.collision_counts(), s.collision_counts()
        assert list(co.icc) == list(cg.icc)
        assert list(co.colltype) == list(cg.colltype)
        assert list(co.nir) == list(cg.nir)
        assert co.ncoll == cg.ncoll
        assert co.pairs_in_dstat == cg.pairs_in_dstat
        bo, bg = o.collision_bins(), s.collision_bins()
        assert np.array_equal(bo["nirbin"], bg["nirbin"])
        for k in ("wrbin", "wxbin", "wybin", "wzbin", "urbin", "thetabin", "wrtanbin"):
            assert np.allclose(bo[k], bg[k], rtol=1e-9, atol=1e-9), k
        # the Vp<i1><i2>.dat records, in the reference's visiting order
        ro, rg = o.contact_records(), s.contact_records()
        assert len(ro) == sum(co.nir) and len(ro) > 3
        assert rg.shape == ro.shape
        assert np.array_equal(ro[:, :2], rg[:, :2])
        assert np.array_equal(ro[:, 2:8], rg[:, 2:8])          # velocities are copied
        assert np.allclose(ro[:, 8:], rg[:, 8:], rtol=0, atol=1e-13)


def test_collision_contact_overflow_matches_oracle(ddns, oracle_mod):
    """5000 pairs placed 1.01 (a + b) apart (inside the 2 % contact shell): more contact
    records than the first buffer holds (4096), so the search and the pair kernel run a
    second time with a wider one; the result is the oracle's."""
    N = 64
    npair = 5000
    p = _params(ddns, N=N, npart=2 * npair, nset=2, radius=[60.0, 90.0], dt=0.01)
    with oracle_mod.Oracle(p) as o, ddns.Solver(p) as s:
        rad = o.scaling()["radius"][:2]
        yp, dm, v0, vn, vp = _collision_inputs(2 * npair, 11, rad, dt=0.01)
        rng = np.random.default_rng(12)
        u = rng.normal(size=(3, npair))
        u /= np.linalg.norm(u, axis=0)
        a, b = yp[6, 0::2], yp[6, 1::2]
        yp[0:3, 1::2] = np.mod(yp[0:3, 0::2] + 1.01 * (a + b) * u, 2 * np.pi)
        o.collision_detect(yp, dm, v0, vn, vp)
        s.collision_detect(yp, dm, v0, vn, vp)
        co, cg = o.collision_counts(), s.collision_counts()
        assert sum(co.nir) > 4096
        assert list(co.nir) == list(cg.nir)
        assert list(co.icc) == list(cg.icc)
        assert co.pairs_in_dstat == cg.pairs_in_dstat
        assert np.array_equal(o.collision_pairs(), s.collision_pairs())
        ro, rg = o.contact_records(), s.contact_records()
        assert rg.shape == ro.shape
        assert np.array_equal(ro[:, :8], rg[:, :8])
        assert np.allclose(ro[:, 8:], rg[:, 8:], rtol=0, atol=1e-13)


def test_collision_candidate_overflow_matches_oracle(ddns, oracle_mod):
    """6000 droplets of 300 um in a 64^3 domain: about 136k pairs inside the statistics
    window (10 collision radii), more than the first list of handed-on pairs holds (65536
    entries at this size), so the search drops what does not fit, reports the count, and
    runs again with a list of that size.  One radius set: the pair kernel sums its 180
    statistics bins per work-group in LDS.  The result is the oracle's."""
    N = 64
    n = 6000
    p = _params(ddns, N=N, npart=n, nset=1, radius=[300.0], dt=0.01)
    with oracle_mod.Oracle(p) as o, ddns.Solver(p) as s:
        rad = o.scaling()["radius"][:1]
        yp, dm, v0, vn, vp = _collision_inputs(n, 11, rad, dt=0.01)
        o.collision_detect(yp, dm, v0, vn, vp)
        s.collision_detect(yp, dm, v0, vn, vp)
        co, cg = o.collision_counts(), s.collision_counts()
        print(f"pairs in the statistics window {co.pairs_in_dstat} (GPU {cg.pairs_in_dstat}), "
              f"collision pairs {co.npairs}", flush=True)
        assert co.pairs_in_dstat > 1 << 16
        assert co.pairs_in_dstat == cg.pairs_in_dstat
        assert list(co.icc) == list(cg.icc) and list(co.nir) == list(cg.nir)
        assert co.npairs > 0 and np.array_equal(o.collision_pairs(), s.collision_pairs())
        bo, bg = o.collision_bins(), s.collision_bins()
        assert np.array_equal(bo["nirbin"], bg["nirbin"])
        for k in ("wrbin", "wxbin", "wybin", "wzbin", "urbin", "thetabin", "wrtanbin"):
            assert np.allclose(bo[k], bg[k], rtol=1e-9, atol=1e-9), k


def test_droplet_steps_match_oracle(ddns, oracle_mod):
    N = 32
    p = _params(ddns, N=N, npart=3000, nset=2, radius=[20.0, 40.0], forcing=1, dt=0.002, rnu=0.02, nonoverlap=1)
    v = random_spectral(N, seed=5, kmax=6)
    with ddns.Solver(p) as s, oracle_mod.Oracle(p) as o:
        s.set_flow(v, v)
        o.set_flow(v, v)
        s.init_droplets(seed=290710)
        o.init_droplets(seed=290710)
        dg, do = s.get_droplets(), o.get_droplets()
        assert np.array_equal(dg["yp"], do["yp"])
        s.step(3)
        o.step(3)
        dg, do = s.get_droplets(), o.get_droplets()
        assert dg["yp"].shape == do["yp"].shape
        assert np.allclose(dg["yp"], do["yp"], rtol=1e-12, atol=1e-12)
        assert np.allclose(dg["vp"], do["vp"], rtol=1e-10, atol=1e-12)
        assert np.allclose(dg["fdp"], do["fdp"], rtol=1e-10, atol=1e-12)
        gv, _ = s.get_flow()
        ov, _ = o.get_flow()
        assert rel_l2(gv, ov) < 1e-11


@pytest.mark.parametrize("hdi", [0, 1])
def test_nonlinear_drag_steps_match_oracle(ddns, oracle_mod, hdi):
    """PART_ADVANCE with the Schiller-Naumann factor (ndrag = 1, partlib.f90:645-651):
    three FP64 steps against the oracle; with HI the relative velocity includes the
    disturbance velocity."""
    N = 32
    kw = dict(N=N, npart=3000, nset=2, radius=[20.0, 40.0], forcing=1, dt=0.002, rnu=0.02, ndrag=1,
              hdi_included=hdi, hdi_trunc=10.0, nonoverlap=hdi)
    v = random_spectral(N, seed=21, kmax=6, amp=3.0)
    with ddns.Solver(_params(ddns, **kw)) as s, oracle_mod.Oracle(_params(ddns, **kw)) as o:
        for x in (s, o):
            x.set_flow(v, v)
            x.init_droplets(seed=2024)
        assert np.array_equal(s.get_droplets()["yp"], o.get_droplets()["yp"])
        s.step(3)
        o.step(3)
        dg, do = s.get_droplets(), o.get_droplets()
        tol = 1e-12 if hdi == 0 else 1e-9
        assert np.allclose(dg["yp"], do["yp"], rtol=tol, atol=tol)
        assert np.allclose(dg["fvp"], do["fvp"], rtol=tol, atol=tol)


def test_overlap_relocation_matches_oracle(ddns, oracle_mod):
    """OVERLAP_DET sweeps that really relocate droplets (partlib.f90:2047-2167,
    RELOCATE 1984-2012): the relocation draws come from one ranff2 stream, so the
    relocated positions and the sweep/relocation counts must be identical."""
    N = 32
    p = _params(ddns, N=N, npart=20000, nset=2, radius=[60.0, 90.0], nonoverlap=1)
    with oracle_mod.Oracle(p) as o, ddns.Solver(p) as s:
        rad = o.scaling()["radius"][:2]
        yp = random_droplets(20000, 13, rad)
        o.set_droplets(yp)
        s.set_droplets(yp)
        o.overlap_fix()
        s.overlap_fix()
        co, cg = o.collision_counts(), s.collision_counts()
        assert co.noverlap_relocated > 5, co.noverlap_relocated
        assert (cg.noverlap_sweeps, cg.noverlap_relocated) == (co.noverlap_sweeps, co.noverlap_relocated)
        dg, do = s.get_droplets(), o.get_droplets()
        assert np.array_equal(dg["yp"], do["yp"])


def test_droplet_steps_with_collisions_match_oracle(ddns, oracle_mod):
    """Three full steps with enough droplets that collisions occur and the
    colliders are relocated (turb.f90:590, istep < nstep_grow)."""
    N = 32
    p = _params(ddns, N=N, npart=40000, nset=2, radius=[60.0, 90.0], forcing=1, dt=0.004, rnu=0.02, nonoverlap=1)
    v = random_spectral(N, seed=9, kmax=6, amp=4.0)
    with ddns.Solver(p) as s, oracle_mod.Oracle(p) as o:
        s.set_flow(v, v)
        o.set_flow(v, v)
        s.init_droplets(seed=4711)
        o.init_droplets(seed=4711)
        assert np.array_equal(s.get_droplets()["yp"], o.get_droplets()["yp"])
        ncoll = 0
        for _ in range(3):
            s.step(1)
            o.step(1)
            cg, co = s.collision_counts(), o.collision_counts()
            assert cg.ncoll == co.ncoll and list(cg.icc) == list(co.icc)
            ncoll += co.ncoll
            dg, do = s.get_droplets(), o.get_droplets()
            assert np.allclose(dg["yp"], do["yp"], rtol=1e-12, atol=1e-12)
        assert ncoll > 0


def test_droplet_steps_with_coalescence_match_oracle(ddns, oracle_mod):
    """COALESCE (partlib.f90:3642-3720) from the first step (nstep_grow = 0 would mean
    the default; 1 makes every step a growth step): merged radii, terminal velocity,
    removal by swapping in the last droplet, in the reference's pair order."""
    N = 32
    p = _params(ddns, N=N, npart=40000, nset=2, radius=[60.0, 90.0], forcing=1, dt=0.004, rnu=0.02,
                nonoverlap=1, nstep_grow=1)
    v = random_spectral(N, seed=9, kmax=6, amp=4.0)
    with ddns.Solver(p) as s, oracle_mod.Oracle(p) as o:
        s.set_flow(v, v)
        o.set_flow(v, v)
        s.init_droplets(seed=4711)
        o.init_droplets(seed=4711)
        removed = 0
        vol0 = float(np.sum(o.get_droplets()["yp"][6] ** 3))
        for _ in range(3):
            s.step(1)
            o.step(1)
            dg, do = s.get_droplets(), o.get_droplets()
            assert dg["yp"].shape == do["yp"].shape
            removed = 40000 - do["yp"].shape[1]
            assert np.array_equal(dg["yp"][7], do["yp"][7])          # same droplets in the same slots
            assert np.allclose(dg["yp"], do["yp"], rtol=1e-13, atol=1e-13)
            assert np.allclose(dg["fvp"], do["fvp"], rtol=1e-12, atol=1e-12)
        assert removed > 0
        # merged pairs keep their volume; only pairs with a halo-image partner (whose
        # partner is not removed, as in the reference) add volume
        assert float(np.sum(do["yp"][6] ** 3)) >= vol0 * (1 - 1e-12)


@pytest.mark.parametrize("prec", [0, 1])
def test_hdi_steps_match_oracle(ddns, oracle_mod, prec):
    """Hydrodynamic interactions (NEI_LIST_PRECOMP, AFUN_PRECOMP, HDI_GMRES,
    partlib.f90:2756-3510) through full steps: the disturbance velocities, the GMRES
    iteration count and the droplet state against the oracle.  FP64: Gram-Schmidt
    order differs, so 1e-9; FP32-mixed: Krylov vectors in float, 1e-4."""
    N = 32
    kw = dict(N=N, npart=4000, nset=2, radius=[20.0, 40.0], forcing=1, dt=0.002, rnu=0.02,
              hdi_included=1, hdi_trunc=10.0)
    v = random_spectral(N, seed=6, kmax=6)
    with ddns.Solver(_params(ddns, precision=prec, **kw)) as s, oracle_mod.Oracle(_params(ddns, **kw)) as o:
        s.set_flow(v, v)
        o.set_flow(v, v)
        s.init_droplets(seed=77)
        o.init_droplets(seed=77)
        for _ in range(2):
            s.step(1)
            o.step(1)
            cg, co = s.collision_counts(), o.collision_counts()
            assert co.gmres_iters > 1
            if prec == 0:
                assert cg.gmres_iters == co.gmres_iters
        dg, do = s.get_droplets(), o.get_droplets()
        pg, po = dg["pertvel"], do["pertvel"]
        scale = np.abs(po).max()
        assert scale > 0
        tol = 1e-9 if prec == 0 else 1e-4
        assert np.abs(pg - po).max() < tol * scale
        assert np.allclose(dg["yp"][:3], do["yp"][:3], rtol=0, atol=1e-9 if prec == 0 else 1e-6)


@pytest.mark.slow
def test_config_128_stochastic_hi_matches_oracle(ddns, oracle_mod):
    """The 128^3 configuration at its own size: parameter.d_128_stoch (stochastic
    forcing, its nu, dt and scales) with 100,000 droplets of 20 / 40 um and HI to 50
    radii.  Two FP64 steps against the oracle."""
    from pathlib import Path
    p = ddns.read_parameter_file(Path(__file__).parent / "golden" / "parameter.d_128_stoch")
    p.npart, p.nset = 100000, 2
    p.radius[0], p.radius[1] = 20.0, 40.0
    p.hdi_included, p.nonoverlap, p.hdi_trunc = 1, 1, 50.0
    N = p.nx
    v = random_spectral(N, seed=12, kmax=10)
    with ddns.Solver(p) as s, oracle_mod.Oracle(p) as o:
        for x in (s, o):
            x.set_flow(v, v)
            x.init_droplets(seed=290710)
        assert np.array_equal(s.get_droplets()["yp"], o.get_droplets()["yp"])
        for _ in range(2):
            s.step(1)
            o.step(1)
            cg, co = s.collision_counts(), o.collision_counts()
            assert cg.gmres_iters == co.gmres_iters
            assert list(cg.icc) == list(co.icc) and cg.ncoll == co.ncoll
        gv, _ = s.get_flow()
        ov, _ = o.get_flow()
        assert rel_l2(gv, ov) < 1e-11
        bg, sg, yg, ivg = s.get_forcing_state()
        bo, so, yo, ivo = o.get_forcing_state()
        assert sg == so and yg == yo and np.array_equal(ivg, ivo)
        dg, do = s.get_droplets(), o.get_droplets()
        assert np.allclose(dg["yp"][:6], do["yp"][:6], rtol=1e-9, atol=1e-10)
        assert np.abs(dg["pertvel"] - do["pertvel"]).max() < 1e-8 * max(1e-30, np.abs(do["pertvel"]).max())


@pytest.mark.parametrize("forcing", [1, 2])
@pytest.mark.parametrize("prec", [0, 1])
def test_split_z_stage_matches_oracle(ddns, oracle_mod, monkeypatch, forcing, prec):
    """The z stage as separate kernels (plain transforms, pointwise update and curl), the
    path the largest grids take, selected here at N = 32 through DDNS_Z_SPLIT."""
    monkeypatch.setenv("DDNS_Z_SPLIT", "1")
    N = 32
    p = _params(ddns, N=N, forcing=forcing, precision=prec, dt=0.002, rnu=0.02)
    v = random_spectral(N, seed=7 + forcing, kmax=6)
    with ddns.Solver(p) as s, oracle_mod.Oracle(_params(ddns, N=N, forcing=forcing, dt=0.002, rnu=0.02)) as o:
        s.set_flow(v, v)
        o.set_flow(v, v)
        o.flow_only(True)
        s.step(10)
        o.step(10)
        gv, go = s.get_flow()
        ov, oo = o.get_flow()
        assert rel_l2(gv, ov) < (1e-11 if prec == 0 else 1e-5)
        assert rel_l2(go, oo) < (1e-10 if prec == 0 else 1e-4)
