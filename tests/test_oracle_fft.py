"""The oracle's 3D FFTs (mpifft3DRC / mpifft3DCR, mpifft.f90:81-198) pinned by
numpy.fft golden vectors in the reference layout, by closed forms and by Parseval."""
from pathlib import Path

import numpy as np
import pytest

GOLD = Path(__file__).parent / "golden" / "fft_golden.npz"


@pytest.fixture(scope="module")
def gold():
    with np.load(GOLD) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("N", [8, 16])
def test_r2c_matches_golden(oracle_mod, gold, N):
    got = oracle_mod.fft_r2c(gold[f"phys_{N}"])
    exp = gold[f"spec_{N}"]
    assert np.abs(got[:N] - exp[:N]).max() < 1e-15 * max(1.0, np.abs(exp).max()) * N


@pytest.mark.parametrize("N", [8, 16])
def test_c2r_matches_golden(oracle_mod, gold, N):
    got = oracle_mod.fft_c2r(gold[f"inv_spec_{N}"])
    exp = gold[f"inv_phys_{N}"]
    assert np.abs(got[:N, :N] - exp[:N, :N]).max() < 1e-12 * np.abs(exp).max()


@pytest.mark.parametrize("mode", [(1, 0, 0), (0, 2, 0), (0, 0, 3), (2, -1, 3), (1, 3, -2)])
def test_single_cosine_mode(oracle_mod, mode):
    """cos(k.x + phi) has exactly two nonzero Fourier coefficients, e^{+-i phi}/2;
    in the half spectrum the stored one is the +k (kx > 0) or, for kx = 0, both
    +k and -k; every other entry must vanish."""
    N = 16
    kx, ky, kz = mode
    phi = 0.3
    x = 2 * np.pi * np.arange(N) / N
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    phys = np.zeros((N + 2, N + 2, N), order="F")
    phys[:N, :N, :] = np.cos(kx * X + ky * Y + kz * Z + phi)
    c = oracle_mod.fft_r2c(phys)
    expect = np.zeros_like(c)
    for s in (1, -1):
        ax, ay, az = s * kx, s * ky, s * kz
        if ax < 0:
            continue
        iz, iy = az % N, ay % N
        val = 0.5 * np.exp(1j * s * phi)
        expect[iz, 2 * ax, iy] += val.real
        expect[iz, 2 * ax + 1, iy] += val.imag
    assert np.abs(c[:N] - expect[:N]).max() < 1e-14


def test_round_trip_and_parseval(oracle_mod):
    N = 16
    rng = np.random.default_rng(5)
    phys = np.zeros((N + 2, N + 2, N), order="F")
    f = rng.standard_normal((N, N, N))
    phys[:N, :N, :] = f
    c = oracle_mod.fft_r2c(phys)
    # Parseval over the half spectrum: kx = 0 and kx = N/2 planes once, others twice
    re, im = c[:N, 0::2, :], c[:N, 1::2, :]
    w = np.full(N // 2 + 1, 2.0)
    w[0] = w[-1] = 1.0
    e_spec = float(np.sum(w[None, :, None] * (re ** 2 + im ** 2)))
    assert abs(e_spec - np.mean(f ** 2)) < 1e-12 * np.mean(f ** 2)
    # the C2R takes the x-Nyquist term from the (zero) padding: clear it, then the
    # round trip reproduces the band-limited part of f exactly
    c2 = c.copy()
    c2[:, N:, :] = 0.0
    back = oracle_mod.fft_c2r(c2)[:N, :N, :]
    nyq = np.fft.irfftn(np.fft.rfftn(f, axes=(1, 2, 0)) * (np.arange(N // 2 + 1) < N // 2)[:, None, None],
                        s=(N, N, N), axes=(1, 2, 0))
    assert np.abs(back - nyq).max() < 1e-12


def test_fft1d_against_numpy(oracle_mod):
    rng = np.random.default_rng(1)
    for n in (2, 4, 8, 32, 64):
        z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        buf = np.empty(2 * n)
        buf[0::2], buf[1::2] = z.real, z.imag
        oracle_mod.lib().orc_fft1d(n, oracle_mod.dp(buf), -1)
        got = buf[0::2] + 1j * buf[1::2]
        assert np.abs(got - np.fft.fft(z)).max() < 1e-13 * n
