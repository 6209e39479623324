"""Two configurations at their own sizes, held to what each states:

- 128^3 with stochastic forcing (parameter.d_128_stoch), 100,000 bidisperse droplets of
  20 / 40 um (a 1:1 mix) and HI to 50 radii: the FP32-mixed path after 10 steps
  against the FP64 oracle's flow (the flow does not see the droplets) and against the
  FP64 GPU path: velocity relative L2 <= 1e-5 and every shell of E(k)
  (postprocessing.f90:151-188) holding more than 1e-12 of the peak within 1e-5;
  the forcing's random state advances identically.
- 1024^3 FP32-mixed with stochastic forcing, 5,000,000 monodisperse 15 um droplets, HI
  to 50 radii and coalescence from the first step (COALESCE, partlib.f90:3642-3825,
  in place of relocation, turb.f90:528, 591), 4 steps, held to properties that do not
  depend on an oracle run: every volume a^3 is a whole multiple of the initial one
  (merged pairs add their volumes), the total volume never drops and grows only by
  pairs whose partner is a periodic image (kept, as in the reference), the droplet count
  drops by the merged pairs, no two droplets overlap after the sweep (brute force on
  the host), the field stays solenoidal and the forcing state advances.  At this
  density monodisperse droplets almost never meet in 4 steps, so 400 pairs are set up
  a hair apart (1e-7 of a diameter beyond contact): the flow's shear closes about half
  of those gaps within the first step, and the merges are then certain.
"""
from pathlib import Path

import numpy as np
import pytest

from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).parent / "golden"


def _shell_err(eg, eo):
    live = eo > 1e-12 * eo.max()
    return float(np.max(np.abs(eg[live] / eo[live] - 1))), int(live.sum())


def test_config_128_fp32_mixed_ten_steps(ddns, oracle_mod):
    base = ddns.read_parameter_file(GOLD / "parameter.d_128_stoch")

    def params(prec, npart):
        p = ddns.read_parameter_file(GOLD / "parameter.d_128_stoch")
        p.precision = prec
        p.npart, p.nset = npart, 2
        p.radius[0], p.radius[1] = 20.0, 40.0
        p.hdi_included, p.nonoverlap, p.hdi_trunc = 1, 1, 50.0
        return p

    assert base.forcing == 2
    with ddns.Solver(params(ddns.FP64, 100_000)) as s:
        s.init_flow(ddns.SPEC_K4_GAUSS, 321, 4.0, 1.0)
        v0, h0 = s.get_flow()
    out = {}
    for prec in (ddns.FP32_MIXED, ddns.FP64):
        with ddns.Solver(params(prec, 100_000)) as s:
            s.set_flow(v0, h0)
            s.init_droplets(seed=290710)
            s.step(10)
            c = s.collision_counts()
            assert c.gmres_iters >= 1
            out[prec] = (s.get_flow()[0], s.spectrum(), s.get_forcing_state())
    with oracle_mod.Oracle(params(ddns.FP64, 0)) as o:
        o.set_flow(v0, h0)
        o.flow_only(True)
        o.step(10)
        vo, eo, fo = o.get_flow()[0], o.spectrum(), o.get_forcing_state()
    v32, e32, f32 = out[ddns.FP32_MIXED]
    v64, e64, f64 = out[ddns.FP64]
    err32, err64 = rel_l2(v32, vo), rel_l2(v64, vo)
    se32, nlive = _shell_err(e32, eo)
    se64, _ = _shell_err(e64, eo)
    print(f"128^3 stochastic, 10 steps: velocity rel L2 FP32-mixed {err32:.3e}, FP64 {err64:.3e} against the oracle; "
          f"E(k) max shell error FP32-mixed {se32:.3e}, FP64 {se64:.3e} over {nlive} shells; "
          f"FP32-mixed against FP64 GPU {rel_l2(v32, v64):.3e}", flush=True)
    assert err64 < 1e-11 and se64 < 1e-10
    assert err32 <= 1e-5 and se32 <= 1e-5 and nlive >= 15
    for f in (f32, f64):                      # iseedf, iyf, ivf of the OU forcing
        assert f[1] == fo[1] and f[2] == fo[2] and np.array_equal(f[3], fo[3])


def _overlapping_pairs(yp):
    """Pairs closer than the sum of their radii, minimum image in the 2 pi box."""
    from scipy.spatial import cKDTree
    box = 2 * np.pi
    pos = np.mod(yp[:3].T, box)
    pos[pos >= box] = 0.0
    a = yp[6]
    tree = cKDTree(pos, boxsize=box)
    pairs = tree.query_pairs(r=2 * a.max(), output_type="ndarray")
    if len(pairs) == 0:
        return 0
    d = pos[pairs[:, 0]] - pos[pairs[:, 1]]
    d -= box * np.round(d / box)
    dist = np.sqrt((d ** 2).sum(axis=1))
    return int(np.count_nonzero(dist < a[pairs[:, 0]] + a[pairs[:, 1]]))


def test_config_1024_stochastic_5M_coalescence(ddns):
    N, n0, r_um = 1024, 5_000_000, 15.0
    # nu for k_max eta ~ 1.5 at u' ~ 1: eta = (nu^3 / eps)^(1/4) ~ 4.8e-3, tau_K ~ 0.028
    p = ddns.default_params(N=N, precision=ddns.FP32_MIXED, dt=0.00025, rnu=0.0008, forcing=2,
                            etk_DNS=0.00476, tauk_DNS=0.0283, npart=n0, nset=1, radius=[r_um],
                            hdi_included=1, hdi_trunc=50.0, nonoverlap=1, nstep_grow=1)
    with ddns.Solver(p) as s:
        s.init_flow(ddns.SPEC_K4_GAUSS, 123, 4.0, 1.0)
        s.init_droplets(seed=290710)
        d0 = s.get_droplets()
        a0 = float(d0["yp"][6][0])
        assert np.all(d0["yp"][6] == a0)
        # droplets 2k + 1 (k < 400) moved next to droplets 2k, 2 a (1 + 1e-7) apart along a
        # random direction; positions wrapped into the box
        yp = np.asfortranarray(d0["yp"].copy())
        rng = np.random.default_rng(5)
        for k in range(400):
            dirn = rng.standard_normal(3)
            dirn /= np.linalg.norm(dirn)
            yp[:3, 2 * k + 1] = np.mod(yp[:3, 2 * k] + 2 * a0 * (1 + 1e-7) * dirn, 2 * np.pi)
        s.set_droplets(yp, d0["fvp"], d0["fdp"], d0["pertvel"])
        d0 = s.get_droplets()
        st0 = s.get_forcing_state()
        n_prev, coll = d0["yp"].shape[1], 0
        assert n_prev == n0
        for k in range(4):
            s.step(1)
            c = s.collision_counts()
            d = s.get_droplets()
            n = d["yp"].shape[1]
            coll += sum(c.icc)
            print(f"step {k + 1}: droplets {n}, collisions {list(c.icc)}, GMRES {c.gmres_iters}, "
                  f"overlap sweeps {c.noverlap_sweeps}", flush=True)
            assert n_prev - n <= sum(c.icc)          # each removal is one merged pair
            n_prev = n
        m = s.monitors(divstat=True)
        st1 = s.get_forcing_state()
        yp = d["yp"]
    units = (yp[6] / a0) ** 3                            # volume in initial droplet volumes
    whole = np.rint(units)
    assert np.all(whole >= 1) and np.abs(units - whole).max() < 1e-9
    grown = float(whole.sum()) - n0                      # volume added: periodic-image partners
    removed = n0 - yp.shape[1]
    print(f"after 4 steps: {removed} droplets removed, {coll} collisions, volume +{grown:.0f} units", flush=True)
    assert coll > 0 and removed > 0
    assert removed >= 0 and 0 <= grown <= coll
    assert _overlapping_pairs(yp) == 0
    assert m.vdiv2 <= 1e-5 * m.vort2
    assert st1[1] != st0[1] or not np.array_equal(st1[3], st0[3])


def test_config_64_plain_fp64_twenty_steps(ddns, oracle_mod):
    """The reference's own CPU-runnable case at its size: 64^3 FP64 (parameter.d_64_det_1:
    deterministic forcing of shells 1-2, nu = 0.026), the k^4 exp(-2 (k/k0)^2) start
    (k0 = 4, u_rms = 1), dt at a CFL number near 0.3, 20 steps, 10,000 monodisperse droplets
    of Stokes number 1 at uniform positions starting with the fluid's velocity, gravity and
    collision detection, no HI.  Against the oracle: the flow to 1e-11, the collision
    counts of every step identical, the droplets to 1e-10, the last step's pair list and
    statistics-window counts identical and its bin sums to 1e-9.  (This dilute case meets
    no collision in 20 steps; some 800 pairs a step fall in the statistics window.)"""
    p = ddns.read_parameter_file(GOLD / "parameter.d_64_det_1")
    p.precision = ddns.FP64
    p.npart, p.nset = 10_000, 1
    p.radius[0] = ddns.stokes_radius_um(p, 1.0)
    p.hdi_included, p.nonoverlap = 0, 0
    p.dt = 0.004
    with ddns.Solver(p) as s:
        s.init_flow(ddns.SPEC_K4_GAUSS, 2024, 4.0, 1.0)
        v0, h0 = s.get_flow()
        cfl = s.monitors().cfl
    assert 0.25 < cfl < 0.4, cfl
    with ddns.Solver(p) as s, oracle_mod.Oracle(p) as o:
        s.set_flow(v0, h0)
        o.set_flow(v0, h0)
        s.init_droplets(seed=290710)
        o.init_droplets(seed=290710)
        total = 0
        for _ in range(20):
            s.step(1)
            o.step(1)
            cg, co = s.collision_counts(), o.collision_counts()
            assert list(cg.icc) == list(co.icc) and cg.ncoll == co.ncoll and cg.npairs == co.npairs
            assert list(cg.colltype) == list(co.colltype) and list(cg.nir) == list(co.nir)
            assert cg.pairs_in_dstat == co.pairs_in_dstat
            total += co.ncoll
        err = rel_l2(s.get_flow()[0], o.get_flow()[0])
        dg, do = s.get_droplets(), o.get_droplets()
        pg, po = s.collision_pairs(), o.collision_pairs()
        bg, bo = s.collision_bins(), o.collision_bins()
    print(f"64^3 FP64, 10k droplets of St 1, 20 steps: velocity rel L2 {err:.2e} against the oracle, "
          f"{total} collisions, droplets max |dx| {np.abs(dg['yp'] - do['yp']).max():.2e}", flush=True)
    assert err < 1e-11
    assert np.abs(dg["yp"] - do["yp"]).max() < 1e-10
    assert np.array_equal(pg, po)
    assert np.array_equal(bg["nirbin"], bo["nirbin"]) and bg["nirbin"].sum() > 0
    for k in ("wrbin", "wxbin", "wybin", "wzbin", "urbin", "thetabin", "wrtanbin"):
        assert np.allclose(bg[k], bo[k], rtol=1e-9, atol=1e-12 * max(1.0, np.abs(bo[k]).max())), k
