"""The tile-shape tuner (the role of the reference's FFTW_MEASURE plans, mpifft.f90:36-56):
at set-up it times half, once and twice the default y-stage tile width, and once and
twice the default work-group size of the fused z stage, and keeps the fastest.
A shape only regroups work, so every candidate must give the same bits; forced
non-default shapes are also held to the oracle."""
import numpy as np
import pytest

from tests.helpers import random_spectral, rel_l2

pytestmark = pytest.mark.gpu


def _candidates(default):
    return [default // 2, default, 2 * default]


def _nt_candidates(default):
    return [default, 2 * default, 4 * default]


@pytest.mark.parametrize("N", [64, 256])
@pytest.mark.parametrize("prec", [0, 1])
def test_tuner_picks_a_timed_candidate(ddns, N, prec):
    with ddns.Solver(ddns.default_params(N=N, precision=prec)) as s:
        t = s.tuning()
    print(f"N={N} prec={prec}: {t}", flush=True)
    for key, cands, ms in (("ky_tc", _candidates(t["ky_tc_default"]), t["ky_ms"]),
                           ("kz_nt", _nt_candidates(t["kz_nt_default"]), t["kz_ms"])):
        chosen = t[key]
        assert chosen in cands, key
        timed = ms > 0
        assert timed[cands.index(t[key + "_default"])], key      # the default always fits
        assert ms[cands.index(chosen)] == ms[timed].min(), key


@pytest.mark.parametrize("prec", [0, 1])
def test_every_tile_width_gives_the_same_flow(ddns, oracle_mod, monkeypatch, prec):
    N = 64
    v = random_spectral(N, seed=21, kmax=10)
    kw = dict(N=N, forcing=1, dt=0.002, rnu=0.02)
    with ddns.Solver(ddns.default_params(precision=prec, **kw)) as s:
        t0 = s.tuning()
    default, nt0 = t0["ky_tc_default"], t0["kz_nt_default"]
    out = {}
    # every y-stage width with the default z stage, then every z-stage size
    shapes = [(tc, nt0) for tc in _candidates(default)] + [(default, nt) for nt in _nt_candidates(nt0)[1:]]
    for tc, nt in shapes:
        monkeypatch.setenv("DDNS_KY_TC", str(tc))
        monkeypatch.setenv("DDNS_KZ_NT", str(nt))
        with ddns.Solver(ddns.default_params(precision=prec, **kw)) as s:
            t = s.tuning()
            if (t["ky_tc"], t["kz_nt"]) != (tc, nt):
                continue                                     # this shape does not fit the tile constraints
            s.set_flow(v, v)
            s.step(5)
            out[(tc, nt)] = s.get_flow()
    assert (default, nt0) in out and len(out) >= 3
    ref = out[(default, nt0)]
    for shape, (u, o) in out.items():
        for a, b in zip(u + o, ref[0] + ref[1]):
            assert np.array_equal(a, b), shape
    with oracle_mod.Oracle(ddns.default_params(**kw)) as o:
        o.set_flow(v, v)
        o.flow_only(True)
        o.step(5)
        ov = o.get_flow()[0]
    for shape in out:
        assert rel_l2(out[shape][0], ov) < (1e-11 if prec == 0 else 1e-5), shape
