"""The oracle's droplet step (partlib.f90) pinned by closed forms, textbook physics
and brute force: Lagrange exactness and sixth-order convergence, the Stokes terminal
velocity and relaxation, a brute-force minimum-image collision search with the cubic
Hermite path written in its basis form, the three collision types, an RDF of one for
uniform droplets, the two-sphere Stokes disturbance problem solved densely, the
no-overlap invariant and coalescence volume."""
import ctypes as C

import numpy as np
import pytest
from scipy.spatial import cKDTree

from tests.helpers import to_spectral

PI2 = 8.0 * np.arctan(1.0)


def _params(**kw):
    from mpi_2d_domain_decomp_turb_particles_amd import _abi
    return _abi.default_params(**kw)


def _yp(pos, rad, vel=None):
    n = pos.shape[1]
    yp = np.zeros((8, n), order="F")
    yp[0:3] = pos
    if vel is not None:
        yp[3:6] = vel
    yp[6] = rad
    yp[7] = np.arange(1, n + 1)
    return yp


# ----------------------------------------------------------------- interpolation
def test_lagrange_weights_exact_for_quintics(oracle_mod):
    lib = oracle_mod.lib()
    dx = 0.3
    node = C.c_int()
    bg = np.zeros(6)
    rng = np.random.default_rng(0)
    for y in rng.uniform(0, 5, 40):
        lib.orc_interp_weights(float(y), dx, C.byref(node), oracle_mod.dp(bg))
        z = y / dx - node.value
        assert 0 <= z < 1
        k = np.arange(-2, 4)                      # nodes node-2 .. node+3
        for d in range(6):
            assert abs(np.dot(bg, k.astype(float) ** d) - z ** d) < 1e-12, d
    lib.orc_interp_weights(3 * dx, dx, C.byref(node), oracle_mod.dp(bg))
    assert np.allclose(bg, [0, 0, 1, 0, 0, 0], atol=1e-15)


def _interp_error(oracle_mod, N):
    x = PI2 * np.arange(N) / N
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    f = [lambda x, y, z: np.sin(2 * y + 0.3), lambda x, y, z: np.cos(3 * z), lambda x, y, z: np.sin(x - 0.7)]
    v = [to_spectral(oracle_mod, g(X, Y, Z), N) for g in f]
    rng = np.random.default_rng(N)
    n = 300
    pos = rng.uniform(0, PI2, size=(3, n))
    p = _params(N=N, npart=n, nset=1, radius=[1.0], forcing=0, dt=1e-4, rnu=0.01)
    with oracle_mod.Oracle(p) as o:
        o.set_flow(v, v)
        o.set_droplets(_yp(pos, o.scaling()["radius"][0]))
        o.step(1)
        vp = o.get_droplets()["vp"]
    exact = np.array([g(pos[0], pos[1], pos[2]) for g in f])
    return np.abs(vp - exact).max()


def test_interpolation_is_sixth_order(oracle_mod):
    """u sampled at random droplets through PREPARE_INTERP/INTERPOLATE: the error of
    the 6-point Lagrange stencil falls by about 2^6 when the grid is refined."""
    e16, e32 = _interp_error(oracle_mod, 16), _interp_error(oracle_mod, 32)
    assert e32 < 5e-4
    assert e16 / e32 > 40


# ----------------------------------------------------------------- advance
def _stokes_scales(p, r_um):
    """Stokes settling in cgs (tau_p = 2 rho_w a^2 / (9 rho_a nu)), mapped to the DNS
    units by the Kolmogorov scales of cloud and DNS."""
    eta_c = (p.viscosity_CLOUD ** 3 / p.ediss_CLOUD) ** 0.25
    tau_c = (p.viscosity_CLOUD / p.ediss_CLOUD) ** 0.5
    a_cm = r_um * 1e-4
    tau_p = 2 * p.rho_water * a_cm ** 2 / (9 * p.rho_air * p.viscosity_CLOUD)
    v_t = tau_p * p.gravity
    v_ratio = (eta_c / tau_c) / (p.etk_DNS / p.tauk_DNS)
    return v_t / v_ratio, tau_p / (tau_c / p.tauk_DNS)


def test_terminal_velocity_is_a_fixed_point(oracle_mod):
    """In still fluid a new droplet starts at the terminal velocity W = tau_p g along
    +x (partlib.f90:568-570) and keeps it: AM4/AB4 coefficients sum to 24/24, so
    dmove = W dt exactly."""
    N, n, dt = 16, 50, 0.002
    p = _params(N=N, npart=n, nset=1, radius=[30.0], forcing=0, dt=dt, rnu=0.01)
    W, _ = _stokes_scales(p, 30.0)
    zero = [np.zeros((N + 2, N + 2, N), order="F") for _ in range(3)]
    rng = np.random.default_rng(1)
    pos = rng.uniform(0, PI2, size=(3, n))
    with oracle_mod.Oracle(p) as o:
        o.set_flow(zero, zero)
        o.set_droplets(_yp(pos, o.scaling()["radius"][0]))
        o.step(3)
        d = o.get_droplets()
    assert np.allclose(d["yp"][3], W, rtol=1e-12)
    assert np.all(d["yp"][4:6] == 0.0)
    moved = (pos[0] + 3 * dt * W) % PI2
    assert np.allclose(d["yp"][0], moved, rtol=0, atol=1e-12)
    assert np.array_equal(d["yp"][1:3], pos[1:3])


def test_relaxation_to_terminal_velocity(oracle_mod):
    """From rest, with the history at rest, V_x(t) = W (1 - exp(-t / tau_p))."""
    N, n, dt, steps = 16, 4, 0.002, 150
    p = _params(N=N, npart=n, nset=1, radius=[40.0], forcing=0, dt=dt, rnu=0.01, new_particle=0)
    W, tau = _stokes_scales(p, 40.0)
    assert 0.05 < tau < 1.0
    zero = [np.zeros((N + 2, N + 2, N), order="F") for _ in range(3)]
    pos = np.array([[0.5, 2.0, 4.0, 5.5], [1.0, 3.0, 2.0, 5.0], [0.3, 1.3, 4.1, 2.9]])
    with oracle_mod.Oracle(p) as o:
        o.set_flow(zero, zero)
        o.set_droplets(_yp(pos, o.scaling()["radius"][0]), np.zeros((3, 3, n), order="F"),
                       np.zeros((3, 4, n), order="F"))
        o.set_new_particle(0)
        o.step(steps)
        v = o.get_droplets()["yp"][3]
    t = steps * dt
    assert np.allclose(v, W * (1 - np.exp(-t / tau)), rtol=2e-3)


# ----------------------------------------------------------------- collisions
def _hermite(x0, x1, v0, v1, dt, t):
    """Cubic Hermite path through (x0, v0 dt) at t = 0 and (x1, v1 dt) at t = 1."""
    h00 = 2 * t ** 3 - 3 * t ** 2 + 1
    h10 = t ** 3 - 2 * t ** 2 + t
    h01 = -2 * t ** 3 + 3 * t ** 2
    h11 = t ** 3 - t ** 2
    return h00 * x0 + h10 * v0 * dt + h01 * x1 + h11 * v1 * dt


def _brute_force_pairs(yp, dmove, v0, v1, dt):
    pos, rad = yp[0:3].T, yp[6]
    tree = cKDTree(pos % PI2, boxsize=PI2)
    reach = 2 * rad.max() + 2 * np.linalg.norm(dmove, axis=0).max()
    out = set()
    for a, b in tree.query_pairs(reach):
        d = pos[a] - pos[b]
        shift = -PI2 * np.round(d / PI2)
        xa0, xb0 = pos[a], pos[b] - shift
        R = rad[a] + rad[b]
        d0 = np.linalg.norm(xa0 - xb0)
        d1 = np.linalg.norm(xa0 + dmove[:, a] - xb0 - dmove[:, b])
        if d0 > R + np.linalg.norm(dmove[:, a]) + np.linalg.norm(dmove[:, b]):
            continue
        if d0 < R and d1 > R:
            continue
        typ = 0
        if d0 > R and d1 <= R:
            typ = 1
        else:
            for k in range(1, 20):
                t = 0.05 * k
                xa = _hermite(xa0, xa0 + dmove[:, a], v0[:, a], v1[:, a], dt, t)
                xb = _hermite(xb0, xb0 + dmove[:, b], v0[:, b], v1[:, b], dt, t)
                dd = np.linalg.norm(xa - xb)
                if d0 > R and dd < R:
                    typ = 2
                    break
                if d1 <= R and dd > R:
                    typ = 3
                    break
        if typ:
            ga, gb = int(yp[7, a]), int(yp[7, b])
            out.add((min(ga, gb), max(ga, gb), typ))
    return out


def test_collision_search_equals_brute_force(oracle_mod):
    """HEADLIST images + COLLISION_DET find exactly the pairs a minimum-image brute
    force finds.  Displacements stay below one grid cell, so the reference's
    adjacent-cell window (partlib.f90:1445) excludes no true pair."""
    N, n, dt = 16, 8000, 0.01
    p = _params(N=N, npart=n, nset=2, radius=[150.0, 300.0], dt=dt)
    with oracle_mod.Oracle(p) as o:
        r = o.scaling()["radius"][:2]
        rng = np.random.default_rng(11)
        pos = rng.uniform(0, PI2, size=(3, n))
        rad = r[rng.integers(0, 2, n)]
        v0 = rng.normal(0, 2.0, size=(3, n))
        v1 = v0 + rng.normal(0, 1.5, size=(3, n))
        dmove = dt / 24 * (9 * v1 + 19 * v0 - 5 * v0 + v0)
        assert np.linalg.norm(dmove, axis=0).max() * 2 + 2 * r.max() < PI2 / N
        yp = _yp(pos, rad, v1)
        vp = np.zeros((3, n), order="F")
        o.collision_detect(yp, np.asfortranarray(dmove), np.asfortranarray(v0), np.asfortranarray(v1), vp)
        got = {tuple(int(x) for x in row) for row in o.collision_pairs()}
    exp = _brute_force_pairs(yp, dmove, v0, v1, dt)
    assert len(exp) > 30
    assert got == exp


def _two_body(oracle_mod, xa, xb, va0, va1, vb0, vb1, dt=0.01, r_um=200.0):
    p = _params(N=16, npart=2, nset=1, radius=[r_um], dt=dt)
    with oracle_mod.Oracle(p) as o:
        a = o.scaling()["radius"][0]
        pos = np.array([xa, xb]).T * a + 3.0
        v0 = np.array([va0, vb0]).T * a / dt
        v1 = np.array([va1, vb1]).T * a / dt
        dm = dt / 24 * (9 * v1 + 19 * v0 - 5 * v0 + v0)
        yp = _yp(np.asfortranarray(pos), a, v1)
        o.collision_detect(yp, np.asfortranarray(dm), np.asfortranarray(v0), np.asfortranarray(v1),
                           np.zeros((3, 2), order="F"))
        return [tuple(int(x) for x in row) for row in o.collision_pairs()]


def test_collision_types_one_two_three(oracle_mod):
    # I: apart at the start, in contact at the end (head-on approach)
    assert _two_body(oracle_mod, [0, 0, 0], [2.5, 0, 0], [1, 0, 0], [1, 0, 0], [0, 0, 0], [0, 0, 0]) == [(1, 2, 1)]
    # II: passes through the partner within the step
    assert _two_body(oracle_mod, [-3, 0, 0], [0, 0.5, 0], [6, 0, 0], [6, 0, 0], [0, 0, 0], [0, 0, 0]) == [(1, 2, 2)]
    # III: in contact at both ends, apart in between: out and back, with v1 = -(15/9) v0
    # so that dmove = dt/24 (9 v1 + 15 v0) = 0 for the histories used here
    assert _two_body(oracle_mod, [0, 0, 0], [-1.8, 0, 0], [1.6, 0, 0], [-1.6 * 15 / 9, 0, 0], [0, 0, 0],
                     [0, 0, 0]) == [(1, 2, 3)]
    # no collision: moving apart
    assert _two_body(oracle_mod, [0, 0, 0], [2.5, 0, 0], [-1, 0, 0], [-1, 0, 0], [0, 0, 0], [0, 0, 0]) == []


@pytest.mark.parametrize("frac", [0.995, 1.005, 1.025, 0.97])
def test_contact_record_of_two_droplets(oracle_mod, frac):
    """One pair at distance frac * (ra + rb), st = 2 %: the contact shell is
    (0.98, 1.02] * rcoll without HI.  Inside it the record is buffer(1:12) of
    partlib.f90:1431-1437 written out by hand: the size classes, both velocities
    fvp(:, 0), the unit vector from the partner to ih1 and the relative velocity along
    it; outside it there is none.  The larger droplet is ih2, so a transposed pair
    or a swapped size class fails."""
    p = _params(N=16, npart=2, nset=2, radius=[150.0, 250.0], dt=0.01)
    with oracle_mod.Oracle(p) as o:
        ra, rb = o.scaling()["radius"][:2]
        u = np.array([2.0, -1.0, 2.0]) / 3.0
        xb = np.array([3.0, 3.1, 2.9])
        xa = xb + frac * (ra + rb) * u
        v0 = np.asfortranarray(np.array([[0.3, -0.2, 0.5], [-0.1, 0.4, 0.2]]).T)
        yp = _yp(np.asfortranarray(np.array([xa, xb]).T), np.array([ra, rb]))
        z = np.zeros((3, 2), order="F")
        o.collision_detect(yp, z, v0, z, z)
        rec = o.contact_records()
        nir = np.array(o.collision_counts().nir).reshape(10, 10, order="F")
    if not 0.98 < frac <= 1.02:
        assert rec.shape == (0, 12) and nir.sum() == 0
        return
    # FINDINDEX orders the two size classes, index1 <= index2 (partlib.f90:1878-1882)
    assert rec.shape == (1, 12) and nir[0, 1] == 1 and nir.sum() == 1
    va, vb = v0[:, 0], v0[:, 1]
    expect = np.concatenate([[1.0, 2.0], va, vb, u, [np.dot(va - vb, u)]])
    assert np.allclose(rec[0], expect, rtol=0, atol=1e-14)


def test_rdf_of_uniform_droplets_is_one(oracle_mod):
    """Pair counts in the 180 bins of width R/20 over (R, 10R] match the ideal-gas
    expectation n(n-1)/2 * shell volume / box volume."""
    N, n = 16, 20000
    p = _params(N=N, npart=n, nset=1, radius=[100.0], dt=0.01)
    with oracle_mod.Oracle(p) as o:
        a = o.scaling()["radius"][0]
        rng = np.random.default_rng(5)
        yp = _yp(rng.uniform(0, PI2, size=(3, n)), a)
        z = np.zeros((3, n), order="F")
        o.collision_detect(yp, z, z, z, z)
        nir = o.collision_bins()["nirbin"][0, 0, :].astype(float)
    R = 2 * a
    edges = R * (1 + 0.05 * np.arange(181))
    expect = n * (n - 1) / 2 * 4 * np.pi / 3 * (edges[1:] ** 3 - edges[:-1] ** 3) / PI2 ** 3
    assert abs(nir.sum() - expect.sum()) < 5 * np.sqrt(expect.sum())
    chi2 = np.sum((nir - expect) ** 2 / expect)
    assert chi2 < 180 + 5 * np.sqrt(2 * 180)


# ----------------------------------------------------------------- hydrodynamic interactions
def _stokes_block(d, a):
    """Disturbance of a Stokes sphere of radius a translating with unit velocity, at
    offset d: 3a/(4r) (I + r r^T/r^2) + a^3/(4r^3) (I - 3 r r^T/r^2)."""
    r = np.linalg.norm(d)
    e = np.outer(d, d) / r ** 2
    return 0.75 * a / r * (np.eye(3) + e) + 0.25 * (a / r) ** 3 * (np.eye(3) - 3 * e)


@pytest.mark.parametrize("sep", [2.5, 4.0, 9.0])
def test_two_sphere_disturbance_matches_dense_solve(oracle_mod, sep):
    p = _params(N=16, npart=3, nset=2, radius=[40.0, 60.0], hdi_included=1, hdi_trunc=50.0)
    with oracle_mod.Oracle(p) as o:
        r = o.scaling()["radius"][:2]
        pos = np.array([[2.0, 2.5, 3.0], [2.0, 2.5, 3.0], [5.0, 1.0, 1.0]]).T
        dirn = np.array([0.6, -0.48, 0.64])
        pos[:, 1] += sep * (r[0] + r[1]) * dirn
        rad = np.array([r[0], r[1], r[0]])
        yp = _yp(np.asfortranarray(pos), rad)
        rng = np.random.default_rng(int(sep * 10))
        V = np.asfortranarray(rng.normal(0, 1, size=(3, 3)))
        U = np.asfortranarray(rng.normal(0, 1, size=(3, 3)))
        st, pv, it, res = o.hdi_solve(yp, V, U, np.zeros((3, 3), order="F"))
    assert st >= 0
    # droplets 1, 2 interact; droplet 3 is beyond every truncation radius
    M = np.zeros((9, 9))
    for i in range(3):
        for j in range(3):
            if i != j:
                d = pos[:, i] - pos[:, j]
                if np.linalg.norm(d) < 50.0 * rad[j]:
                    M[3 * i:3 * i + 3, 3 * j:3 * j + 3] = _stokes_block(d, rad[j])
    b = M @ (V - U).T.reshape(-1)
    x = np.linalg.solve(np.eye(9) + M, b)
    assert np.abs(M[6:, :]).max() == 0 and np.abs(M[:, 6:]).max() == 0
    assert np.allclose(pv.T.reshape(-1), x, rtol=0, atol=1e-6 * np.abs(x).max())
    assert np.abs(x).max() > 1e-3


# ----------------------------------------------------------------- overlap and coalescence
def test_overlap_det_leaves_no_overlap(oracle_mod):
    N, n = 16, 6000
    p = _params(N=N, npart=n, nset=2, radius=[150.0, 250.0], nonoverlap=1)
    with oracle_mod.Oracle(p) as o:
        r = o.scaling()["radius"][:2]
        rng = np.random.default_rng(3)
        pos = rng.uniform(0, PI2, size=(3, n))
        rad = r[rng.integers(0, 2, n)]
        yp = _yp(pos, rad)
        tree = cKDTree(pos.T, boxsize=PI2)
        before = sum(1 for a, b in tree.query_pairs(2 * r.max())
                     if np.linalg.norm((pos[:, a] - pos[:, b] + PI2 / 2) % PI2 - PI2 / 2) <= rad[a] + rad[b])
        o.set_droplets(yp)
        o.overlap_fix()
        cc = o.collision_counts()
        d = o.get_droplets()["yp"]
    assert before > 10 and cc.noverlap_relocated >= before
    q = d[0:3] % PI2
    tree = cKDTree(q.T, boxsize=PI2)
    for a, b in tree.query_pairs(2 * r.max()):
        sep = np.linalg.norm((q[:, a] - q[:, b] + PI2 / 2) % PI2 - PI2 / 2)
        assert sep > d[6, a] + d[6, b]
    assert np.array_equal(np.sort(d[7]), np.arange(1, n + 1))   # droplets relocated, none lost


def test_coalescence_conserves_volume(oracle_mod):
    """Two droplets collide head-on inside the box (no halo images): one droplet of
    volume-equivalent radius remains, falling at its terminal velocity."""
    N, dt = 16, 0.01
    p = _params(N=N, npart=2, nset=2, radius=[200.0, 300.0], forcing=0, dt=dt, rnu=0.01,
                nonoverlap=1, nstep_grow=1, new_particle=0)
    zero = [np.zeros((N + 2, N + 2, N), order="F") for _ in range(3)]
    with oracle_mod.Oracle(p) as o:
        sc = o.scaling()
        a1, a2 = sc["radius"][:2]
        W1, W2 = sc["Sv_const"] * a1 ** 2, sc["Sv_const"] * a2 ** 2
        # the larger droplet sits below the smaller one and falls faster along +x
        gap = (W2 - W1) * dt * 0.5
        pos = np.array([[3.0 + a1 + a2 + gap, 3.0, 3.0], [3.0, 3.0, 3.0]]).T
        vel = np.array([[W1, 0, 0], [W2, 0, 0]]).T
        fvp = np.zeros((3, 3, 2), order="F")
        for t in range(3):
            fvp[:, t, :] = vel
        o.set_flow(zero, zero)
        o.set_droplets(_yp(np.asfortranarray(pos), np.array([a1, a2]), vel), fvp, np.zeros((3, 4, 2), order="F"))
        o.set_new_particle(0)
        o.step(1)
        d = o.get_droplets()["yp"]
    assert d.shape[1] == 1
    a = (a1 ** 3 + a2 ** 3) ** (1 / 3)
    assert abs(d[6, 0] - a) < 1e-15 * a
    assert abs(d[3, 0] - sc["Sv_const"] * a * a) < 1e-12 * abs(d[3, 0])
