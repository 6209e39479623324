"""Batched realisations (ddns_batch): every flow stage one launch over all realisations,
each realisation's droplet block as a single handle runs it.  A batch must give each
realisation exactly what a handle of its own gives on the same inputs:

- flow alone, four realisations with deterministic, stochastic (two seeds) and no forcing,
  FP64 and FP32-mixed: the same bits after 6 steps, and the same forcing random state;
- droplets with HI, collisions and relocation, three realisations with uneven droplet
  counts and a droplet size each, FP64: the same flow, droplets, GMRES iterations and
  collision counts after 3 steps, with the HI solves batched and solved one by one;
- the 48-realisation 64^3 configuration (uneven counts 5k-200k, Stokes numbers 0.2-3,
  mixed forcing, HI, FP32-mixed): 2 steps run, and three of its realisations, re-run as
  handles of their own, agree bit for bit.
"""
import numpy as np
import pytest

from tests.helpers import random_spectral

pytestmark = pytest.mark.gpu


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_flow_equals_single_handles(ddns, prec):
    N = 32
    kinds = [(1, 0), (2, 232300), (2, 98765), (0, 0)]       # forcing kind, iseedf
    params = []
    for forcing, seed in kinds:
        p = ddns.default_params(N=N, precision=prec, forcing=forcing, dt=0.002, rnu=0.02)
        if seed:
            p.iseedf = seed
        params.append(p)
    fields = [random_spectral(N, seed=40 + r, kmax=6) for r in range(len(kinds))]
    steps = 6
    single = []
    for p, v in zip(params, fields):
        with ddns.Solver(p) as s:
            s.set_flow(v, v)
            s.step(steps)
            single.append((s.get_flow(), s.get_forcing_state(), s.spectrum()))
    with ddns.Batch(params) as b:
        for r, v in enumerate(fields):
            b.member(r).set_flow(v, v)
        b.step(2)
        b.step(steps - 2)                      # two calls: the graphs are replayed
        t = b.timing()
        assert t.nsteps == steps - 2 and t.flow_ms > 0
        for r in range(len(kinds)):
            m = b.member(r)
            (u, o), fs, ek = m.get_flow(), m.get_forcing_state(), m.spectrum()
            (u1, o1), fs1, ek1 = single[r]
            print(f"realisation {r}: forcing {kinds[r][0]}, max |u - u_single| "
                  f"{max(float(np.max(np.abs(x - y))) for x, y in zip(u, u1)):.3e}", flush=True)
            assert _same(u, u1) and _same(o, o1), f"realisation {r} differs"
            assert np.array_equal(ek, ek1)
            assert fs[1] == fs1[1] and fs[2] == fs1[2] and np.array_equal(fs[3], fs1[3])
            assert np.array_equal(fs[0], fs1[0])


@pytest.mark.parametrize("merged", ["1", "0"])
def test_batch_droplets_hi_equal_single_handles(ddns, monkeypatch, merged):
    """merged = 1: the HI systems of all realisations in one batched GMRES; 0: each
    realisation solves its own inside its droplet block (DDNS_BATCH_HI=0)."""
    monkeypatch.setenv("DDNS_BATCH_HI", merged)
    N = 64
    spec = [(3000, 30.0, 1), (9000, 20.0, 2), (1500, 45.0, 1)]   # droplets, radius (um), forcing
    params = []
    for r, (n, a, forcing) in enumerate(spec):
        p = ddns.default_params(N=N, precision=ddns.FP64, forcing=forcing, npart=n, radius=[a],
                                hdi_included=1, hdi_trunc=20.0, iseedf=232300 + r, relocate_seed=r)
        params.append(p)
    steps = 3
    single = []
    for r, p in enumerate(params):
        with ddns.Solver(p) as s:
            s.init_flow(ddns.SPEC_K4_GAUSS, 500 + r, 4.0, 1.0)
            s.init_droplets(seed=290710 + r)
            s.step(steps)
            single.append((s.get_flow(), s.get_droplets(), s.collision_counts(), s.collision_pairs()))
    with ddns.Batch(params) as b:
        for r in range(len(spec)):
            m = b.member(r)
            m.init_flow(ddns.SPEC_K4_GAUSS, 500 + r, 4.0, 1.0)
            m.init_droplets(seed=290710 + r)
        b.step(steps)
        for r in range(len(spec)):
            m = b.member(r)
            (u, o), d, c, pairs = m.get_flow(), m.get_droplets(), m.collision_counts(), m.collision_pairs()
            (u1, o1), d1, c1, pairs1 = single[r]
            print(f"realisation {r}: {spec[r][0]} droplets, GMRES {c.gmres_iters} / {c1.gmres_iters} iterations, "
                  f"{c.npairs} pairs", flush=True)
            assert _same(u, u1) and _same(o, o1), f"flow of realisation {r} differs"
            for k in ("yp", "fvp", "fdp", "pertvel", "vp"):
                assert np.array_equal(d[k], d1[k]), f"{k} of realisation {r} differs"
            assert c.gmres_iters == c1.gmres_iters and c.gmres_iters >= 1
            assert list(c.icc) == list(c1.icc) and c.npairs == c1.npairs
            assert np.array_equal(pairs, pairs1)


def test_batch_48_realisations_64(ddns):
    params = ddns.batched_realisations(R=48, N=64, precision=ddns.FP32_MIXED)
    counts = [p.npart for p in params]
    assert min(counts) >= 5000 and max(counts) <= 200000
    assert {p.forcing for p in params} == {1, 2}
    steps = 2
    check = [0, 23, 47]
    with ddns.Batch(params) as b:
        for r, p in enumerate(params):
            m = b.member(r)
            m.init_flow(ddns.SPEC_K4_GAUSS, p.iseed_m0, 4.0, 1.0)
            m.init_droplets(seed=290710 + r)
        b.step(steps)
        t = b.timing()
        got = {}
        iters = []
        for r in range(len(params)):
            c = b.member(r).collision_counts()
            iters.append(c.gmres_iters)
            if r in check:
                got[r] = (b.member(r).get_flow()[0], b.member(r).get_droplets())
        print(f"48 x 64^3, {sum(counts)} droplets: {t.total_ms / steps:.2f} ms per step "
              f"(flow {t.flow_ms / steps:.2f} ms); GMRES iterations {min(iters)}..{max(iters)}", flush=True)
        assert min(iters) >= 1
    for r in check:
        p = params[r]
        with ddns.Solver(p) as s:
            s.init_flow(ddns.SPEC_K4_GAUSS, p.iseed_m0, 4.0, 1.0)
            s.init_droplets(seed=290710 + r)
            s.step(steps)
            u1, d1 = s.get_flow()[0], s.get_droplets()
        u, d = got[r]
        assert _same(u, u1), f"flow of realisation {r} differs"
        assert np.array_equal(d["yp"], d1["yp"]) and np.array_equal(d["pertvel"], d1["pertvel"])


@pytest.mark.parametrize("prec", [0, 1])
def test_initial_field_is_reproducible(ddns, prec):
    """initial_flow on the device twice with one seed: the same bits (its shell sums run in
    a fixed order)."""
    out = []
    for _ in range(2):
        with ddns.Solver(ddns.default_params(N=64, precision=prec)) as s:
            s.init_flow(ddns.SPEC_K4_GAUSS, 77, 4.0, 1.0)
            out.append(s.get_flow()[0])
    assert _same(out[0], out[1])


def test_batch_mixed_members_equal_single_handles(ddns):
    """A batch whose realisations differ in what their droplet blocks do: no droplets, droplets
    without HI (collisions and relocation only), HI with the usual GMRES limit, and HI with
    another limit (solved on its own, outside the batched GMRES).  Each gives what a handle
    of its own gives."""
    N = 64
    base = dict(N=N, precision=ddns.FP64, forcing=1)
    params = [
        ddns.default_params(**base),
        ddns.default_params(**base, npart=4000, radius=[50.0], nonoverlap=1, relocate_seed=1),
        ddns.default_params(**base, npart=3000, radius=[30.0], hdi_included=1, hdi_trunc=20.0, relocate_seed=2),
        ddns.default_params(**base, npart=2000, radius=[35.0], hdi_included=1, hdi_trunc=20.0, relocate_seed=3,
                            gmres_maxiter=30),
    ]
    steps = 3
    single = []
    for r, p in enumerate(params):
        with ddns.Solver(p) as s:
            s.init_flow(ddns.SPEC_K4_GAUSS, 700 + r, 4.0, 1.0)
            if p.npart:
                s.init_droplets(seed=1000 + r)
            s.step(steps)
            single.append((s.get_flow()[0], s.get_droplets() if p.npart else None, s.collision_counts()))
    with ddns.Batch(params) as b:
        for r, p in enumerate(params):
            m = b.member(r)
            m.init_flow(ddns.SPEC_K4_GAUSS, 700 + r, 4.0, 1.0)
            if p.npart:
                m.init_droplets(seed=1000 + r)
        b.step(steps)
        for r, p in enumerate(params):
            m = b.member(r)
            u1, d1, c1 = single[r]
            assert _same(m.get_flow()[0], u1), f"flow of realisation {r} differs"
            if p.npart:
                d, c = m.get_droplets(), m.collision_counts()
                for k in ("yp", "fvp", "fdp", "pertvel"):
                    assert np.array_equal(d[k], d1[k]), f"{k} of realisation {r} differs"
                assert c.gmres_iters == c1.gmres_iters and list(c.icc) == list(c1.icc)
                if p.hdi_included:
                    assert c.gmres_iters >= 1
