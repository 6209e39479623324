"""The oracle's initial field (initial_flow, flowlib.f90:53-144, with a counter-based
SplitMix64 generator in place of IFPORT rand), pinned by the generator's published
output stream and by what the construction fixes exactly: a solenoidal, Hermitian,
dealiased field whose shell energies are the requested spectrum."""
import numpy as np
import pytest

from mpi_2d_domain_decomp_turb_particles_amd import _abi

N = 32


def _wavenumbers(N):
    L = N + 2
    i = np.arange(L)
    kz = np.where(i < N, np.where(i <= N // 2, i, i - N), 0).astype(float)
    kx = (np.arange(L) // 2).astype(float)
    k3 = np.arange(N)
    ky = np.where(k3 <= N // 2, k3, k3 - N).astype(float)
    return np.meshgrid(kz, kx, ky, indexing="ij")


def _complex(v):
    """reference layout v(kz, kx re/im, ky) -> complex c[kz, kx, ky], kz < N, kx <= N/2"""
    return v[:N, 0::2, :] + 1j * v[:N, 1::2, :]


def _init(oracle_mod, kind, seed=123, k0=4.0, urms=1.0, **kw):
    p = _abi.default_params(N=N, **kw)
    with oracle_mod.Oracle(p) as o:
        o.init_flow(kind, seed, k0, urms)
        return o.get_flow(), int(N / 2 - 1.5)                     # nek, turb.f90:303-304


def test_splitmix64_published_stream(oracle_mod):
    lib = oracle_mod.lib()
    assert [lib.orc_splitmix64(0, i) for i in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                                          0x06C45D188009454F]


def _shells(c):
    KZ, KX, KY = np.meshgrid(np.where(np.arange(N) <= N // 2, np.arange(N), np.arange(N) - N),
                             np.arange(N // 2 + 1),
                             np.where(np.arange(N) <= N // 2, np.arange(N), np.arange(N) - N), indexing="ij")
    k2 = (KX ** 2 + KY ** 2 + KZ ** 2).astype(float)
    k2[0, 0, 0] = 1e-5
    sh = (np.sqrt(k2) + 0.5).astype(int)
    w = np.where(KX == 0, 0.5, 1.0)
    return KZ, KX, KY, sh, w


@pytest.mark.parametrize("kind", [_abi.SPEC_K4_GAUSS, _abi.SPEC_KOLMOGOROV])
def test_initial_field_is_solenoidal_hermitian_and_on_spectrum(oracle_mod, kind):
    (v, o), nek = _init(oracle_mod, kind)
    for a, b in zip(v, o):
        assert np.array_equal(a, b)                               # history = velocity
    c = [_complex(x) for x in v]
    KZ, KX, KY, sh, w = _shells(c)
    # solenoidal
    div = KX * c[0] + KY * c[1] + KZ * c[2]
    amp = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2) * np.sqrt(sum(np.abs(x) ** 2 for x in c))
    assert np.abs(div).max() <= 1e-13 * amp.max()
    # Nyquist planes, the mean and everything above nek are zero
    for x in c:
        assert np.all(x[N // 2] == 0) and np.all(x[:, :, N // 2] == 0) and np.all(x[:, N // 2] == 0)
        assert x[0, 0, 0] == 0
        assert np.all(x[sh > nek] == 0)
    # Hermitian kx = 0 plane: c(-kz, -ky) = conj c(kz, ky)
    for x in c:
        p = x[:, 0, :]
        m = np.conj(p[(-np.arange(N)) % N][:, (-np.arange(N)) % N])
        assert np.array_equal(p, m)
    e = sum(np.abs(x) ** 2 for x in c) * w
    E = np.array([e[sh == s].sum() for s in range(1, nek + 1)])
    if kind == _abi.SPEC_K4_GAUSS:
        q = np.arange(1, nek + 1) / 4.0
        target = q ** 4 * np.exp(-2 * q * q)
        target *= 1.5 / target.sum()                              # 3/2 u_rms^2, u_rms = 1
        assert np.allclose(E, target, rtol=1e-12, atol=0)
    else:
        # shell 1 holds force(1); above it the shell energy per unit shell volume, per
        # mode count, falls as k^(-5/3) (flowlib.f90:113-123)
        cnt = np.array([w[sh == s].sum() for s in range(1, nek + 1)])
        s = np.arange(1, nek + 1)
        vol = 4 / 3 * np.pi * ((s + 0.5) ** 3 - (s - 0.5) ** 3)
        force1 = _abi.default_params(N=N).force[0]
        assert abs(E[0] - force1) <= 1e-12 * force1
        ratio = (E / (cnt / vol)) * s ** (2 * 0.833333)         # the reference's exponent
        assert np.allclose(ratio, ratio[0], rtol=1e-12)


def test_initial_field_depends_on_the_seed_only(oracle_mod):
    (a, _), _ = _init(oracle_mod, _abi.SPEC_K4_GAUSS, seed=5)
    (b, _), _ = _init(oracle_mod, _abi.SPEC_K4_GAUSS, seed=5)
    (c, _), _ = _init(oracle_mod, _abi.SPEC_K4_GAUSS, seed=6)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not np.allclose(a[0], c[0])


def test_initial_field_at_rest(oracle_mod):
    (v, o), _ = _init(oracle_mod, -1)
    assert all(np.all(x == 0) for x in v + o)
