"""The parameter sets of the batched 48 x 64^3 configuration (host side, no GPU): droplet
counts drawn from [5000, 200000], deterministic and stochastic forcing alternating,
distinct seeds, one droplet size per realisation whose Stokes number tau_p / tau_k --
recomputed here from the radius with tau_p = 2 rho_w a^2 / (9 rho_a nu) and
tau_k = (nu / eps)^(1/2) -- runs evenly over [0.2, 3], and an HI truncation that keeps the
mean neighbour count of a uniform droplet cloud at or below 24."""
import numpy as np

import mpi_2d_domain_decomp_turb_particles_amd as dd


def test_batched_realisations_shape():
    ps = dd.batched_realisations(R=48, N=64, precision=dd.FP32_MIXED)
    assert len(ps) == 48
    counts = np.array([p.npart for p in ps])
    assert counts.min() >= 5000 and counts.max() <= 200000 and len(set(counts)) > 40
    assert [p.forcing for p in ps[:4]] == [1, 2, 1, 2]
    assert len({p.iseedf for p in ps}) == 48 and len({p.iseed_m0 for p in ps}) == 48
    assert len({p.relocate_seed for p in ps}) == 48
    assert all(p.nx == p.ny == p.nz == 64 and p.nset == 1 and p.precision == dd.FP32_MIXED for p in ps)
    assert len({(p.dt, p.rnu) for p in ps}) == 1        # one batch: shared dt and viscosity


def test_batched_realisations_stokes_numbers():
    ps = dd.batched_realisations(R=48, N=64)
    p = ps[0]
    tau_k = np.sqrt(p.viscosity_CLOUD / p.ediss_CLOUD)
    st = []
    for q in ps:
        a = q.radius[0] * 1e-4                                            # um -> cm
        tau_p = 2.0 * q.rho_water * a * a / (9.0 * q.rho_air * q.viscosity_CLOUD)
        st.append(tau_p / tau_k)
    assert np.allclose(st, np.linspace(0.2, 3.0, 48), rtol=1e-12)
    # about 17.7 um at St 0.2 and 68.7 um at St 3 for the reference's cloud constants
    assert abs(ps[0].radius[0] - 17.75) < 0.05 and abs(ps[-1].radius[0] - 68.75) < 0.05


def test_batched_realisations_hi_truncation():
    ps = dd.batched_realisations(R=48, N=64)
    vol = (2 * np.pi) ** 3
    for p in ps:
        assert p.hdi_included == 1 and p.nonoverlap == 1 and 0 < p.hdi_trunc <= 50.0
        etk_cloud = (p.viscosity_CLOUD ** 3 / p.ediss_CLOUD) ** 0.25
        a = p.radius[0] * (p.etk_DNS / etk_cloud) / 1e4                  # code units
        mean_nei = p.npart / vol * 4.0 / 3.0 * np.pi * (p.hdi_trunc * a) ** 3
        assert mean_nei <= 24.0 * (1 + 1e-9)
        if p.hdi_trunc < 50.0:
            assert abs(mean_nei - 24.0) < 1e-6
