"""The headline workload at its full size against the FP64 oracle: 256^3, deterministic
forcing (parameter.d_256_det's nu, dt and Kolmogorov scales), 1,000,000 droplets with
log-normal radii (median 25 um, sigma_ln 0.4, clipped to 5-60 um), hydrodynamic
interactions to 50 radii, collision detection, statistics and relocation.

- FP64: two full steps (turb.f90:492-707) from the same state as the oracle.  The
  velocity to 1e-11 relative L2; the GMRES iteration counts (partlib.f90:2795-3005) and
  every integer collision output (pair list, type counts, colltype, contact counts nir,
  removal-list size; partlib.f90:1166-1846) identical; the disturbance velocities to
  1e-8 of their largest value (Gram-Schmidt order differs), positions to 1e-9.
- FP32-mixed: ten steps; the flow does not see the droplets, so the oracle runs the
  flow alone.  Velocity relative L2 <= 1e-5 (the precision budget of the FP32-mixed
  mode) and the per-shell energy spectrum (postprocessing.f90:151-188) within 1e-5 on
  every shell holding more than 1e-12 of the peak shell energy.
"""
import numpy as np
import pytest

from tests.helpers import rel_l2

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

N, NPART = 256, 1_000_000


def _params(ddns, prec, npart=NPART):
    # parameter.d_256_det: rnu 0.002, dt 0.001, eta/tau_K 0.01392/0.09694, force 1.242477/0.391356
    return ddns.default_params(
        N=N, precision=prec, dt=0.001, rnu=0.002, forcing=1, force=(1.242477, 0.391356),
        etk_DNS=0.01392, tauk_DNS=0.09694, npart=npart, nset=10,
        radius=[5.0, 10.0, 15.0, 20.0, 25.0, 30.0, 35.0, 40.0, 50.0, 60.0],
        hdi_included=1, hdi_trunc=50.0, nonoverlap=1,
    )


def _init(x, ddns):
    x.init_droplets(seed=290710, radii=ddns.RADII_LOGNORMAL, median_um=25.0, sigma_ln=0.4, rmin_um=5.0,
                    rmax_um=60.0)


def test_headline_fp64_two_steps_match_oracle(ddns, oracle_mod):
    with ddns.Solver(_params(ddns, ddns.FP64)) as s, oracle_mod.Oracle(_params(ddns, ddns.FP64)) as o:
        s.init_flow(ddns.SPEC_K4_GAUSS, 123, 4.0, 1.0)
        v, h = s.get_flow()
        o.set_flow(v, h)
        _init(s, ddns)
        _init(o, ddns)
        assert np.array_equal(s.get_droplets()["yp"], o.get_droplets()["yp"])        # PART_INIT + OVERLAP_DET
        pairs = 0
        for k in range(2):
            s.step(1)
            o.step(1)
            cg, co = s.collision_counts(), o.collision_counts()
            print(f"step {k + 1}: GMRES {cg.gmres_iters}/{co.gmres_iters}, icc {list(cg.icc)}/{list(co.icc)}, "
                  f"nir {sum(cg.nir)}/{sum(co.nir)}, ncoll {cg.ncoll}/{co.ncoll}", flush=True)
            assert co.gmres_iters > 1 and cg.gmres_iters == co.gmres_iters
            assert list(cg.icc) == list(co.icc)
            assert list(cg.colltype) == list(co.colltype)
            assert list(cg.nir) == list(co.nir)
            assert cg.ncoll == co.ncoll and cg.npairs == co.npairs
            assert cg.pairs_in_dstat == co.pairs_in_dstat
            pg, po = s.collision_pairs(), o.collision_pairs()
            assert np.array_equal(pg, po)
            pairs += len(po)
        assert pairs > 0
        gv, _ = s.get_flow()
        ov, _ = o.get_flow()
        assert rel_l2(gv, ov) < 1e-11
        dg, do = s.get_droplets(), o.get_droplets()
        assert dg["yp"].shape == do["yp"].shape
        assert np.array_equal(dg["yp"][7], do["yp"][7])
        assert np.allclose(dg["yp"][:6], do["yp"][:6], rtol=1e-9, atol=1e-9)
        pv = np.abs(do["pertvel"]).max()
        assert pv > 0
        assert np.abs(dg["pertvel"] - do["pertvel"]).max() < 1e-8 * pv


def test_headline_fp32_mixed_ten_steps_vs_oracle(ddns, oracle_mod):
    with ddns.Solver(_params(ddns, ddns.FP32_MIXED)) as s, \
            oracle_mod.Oracle(_params(ddns, ddns.FP64, npart=0)) as o:
        s.init_flow(ddns.SPEC_K4_GAUSS, 123, 4.0, 1.0)
        v, h = s.get_flow()
        o.set_flow(v, h)
        _init(s, ddns)
        s.step(10)
        o.flow_only(True)
        o.step(10)
        gv, _ = s.get_flow()
        ov, _ = o.get_flow()
        err = rel_l2(gv, ov)
        eg, eo = s.spectrum(), o.spectrum()
    live = eo > 1e-12 * eo.max()
    shell_err = np.abs(eg[live] / eo[live] - 1)
    print(f"FP32-mixed after 10 steps: velocity rel L2 {err:.3e}; E(k) max rel error {shell_err.max():.3e} over "
          f"{live.sum()} shells (worst shell {np.flatnonzero(live)[shell_err.argmax()] + 1})", flush=True)
    assert err <= 1e-5
    assert live.sum() > 20
    assert shell_err.max() <= 1e-5
