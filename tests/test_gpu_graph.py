"""The flow step replayed from captured HIP graphs against the same kernels launched one
by one (DDNS_GRAPH=0): the same bits, flow alone (where a call waits for the device
only once, at its end) and with droplets, HI, collisions and relocation beside it."""
import numpy as np
import pytest

from tests.helpers import random_spectral

pytestmark = pytest.mark.gpu


def _run(ddns, monkeypatch, graph, p, v, steps, droplets):
    monkeypatch.setenv("DDNS_GRAPH", "1" if graph else "0")
    with ddns.Solver(p) as s:
        s.set_flow(v, v)
        if droplets:
            s.init_droplets(seed=4711)
        pairs = []
        for _ in range(steps):
            s.step(1)
            if droplets:
                pairs.append(s.collision_pairs())
        t = s.timing()
        out = (s.get_flow(), s.get_droplets() if droplets else None, pairs, t)
        if not droplets:
            s.step(steps)                      # several steps in one call: one wait at the end
            out = out + (s.get_flow(), s.timing())
    return out


@pytest.mark.parametrize("prec", [0, 1])
def test_flow_graph_equals_eager(ddns, monkeypatch, prec):
    N = 32
    p = ddns.default_params(N=N, precision=prec, forcing=2, dt=0.002, rnu=0.02)
    v = random_spectral(N, seed=31, kmax=6)
    g = _run(ddns, monkeypatch, True, p, v, 6, False)
    e = _run(ddns, monkeypatch, False, p, v, 6, False)
    for a, b in zip(g[0][0] + g[0][1], e[0][0] + e[0][1]):
        assert np.array_equal(a, b)
    for a, b in zip(g[4][0] + g[4][1], e[4][0] + e[4][1]):
        assert np.array_equal(a, b)
    tg, te = g[5], e[5]
    assert tg.nsteps == 6 and tg.flow_ms > 0 and tg.total_ms > 0
    assert all(x == 0 for x in tg.flow_kernel_ms)          # one wait per call: no per-kernel split
    assert all(x > 0 for x in te.flow_kernel_ms[:3])


def test_step_with_droplets_graph_equals_eager(ddns, monkeypatch):
    """FP64, 40,000 droplets of 60 / 90 um, HI to 10 radii, collisions with relocation."""
    N = 32
    p = ddns.default_params(N=N, npart=40000, nset=2, radius=[60.0, 90.0], forcing=1, dt=0.004, rnu=0.02,
                            nonoverlap=1, hdi_included=1, hdi_trunc=10.0)
    v = random_spectral(N, seed=9, kmax=6, amp=4.0)
    g = _run(ddns, monkeypatch, True, p, v, 4, True)
    e = _run(ddns, monkeypatch, False, p, v, 4, True)
    for a, b in zip(g[0][0] + g[0][1], e[0][0] + e[0][1]):
        assert np.array_equal(a, b)
    for k in ("yp", "fvp", "fdp", "pertvel"):
        assert np.array_equal(g[1][k], e[1][k]), k
    assert sum(len(x) for x in e[2]) > 0
    for a, b in zip(g[2], e[2]):
        assert np.array_equal(a, b)
    assert all(x > 0 for x in g[3].flow_kernel_ms[:5])     # the graph's timing events


@pytest.mark.parametrize("graph", [0, 1])
def test_z_stage_gated_on_pair_launch_equals_default(ddns, monkeypatch, graph):
    """DDNS_Z_GATE=1 (the default) queues the flow's second half in two pieces, the z
    stage after the droplet block's collision-pair launch; DDNS_Z_GATE=0 queues it in
    one.  Only the order of launches changes."""
    N = 32
    p = ddns.default_params(N=N, npart=40000, nset=2, radius=[60.0, 90.0], forcing=1, dt=0.004, rnu=0.02,
                            nonoverlap=1, hdi_included=1, hdi_trunc=10.0)
    v = random_spectral(N, seed=9, kmax=6, amp=4.0)
    monkeypatch.setenv("DDNS_Z_GATE", "1")
    g = _run(ddns, monkeypatch, graph, p, v, 4, True)
    monkeypatch.setenv("DDNS_Z_GATE", "0")
    e = _run(ddns, monkeypatch, graph, p, v, 4, True)
    for a, b in zip(g[0][0] + g[0][1], e[0][0] + e[0][1]):
        assert np.array_equal(a, b)
    for k in ("yp", "fvp", "fdp", "pertvel"):
        assert np.array_equal(g[1][k], e[1][k]), k
    for a, b in zip(g[2], e[2]):
        assert np.array_equal(a, b)
    assert all(x > 0 for x in g[3].flow_kernel_ms[:5])
