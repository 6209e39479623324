"""Shared helpers for the parity tests: synthetic inputs and field comparisons."""
from __future__ import annotations

import numpy as np


def random_spectral(N, seed=0, kmax=None, amp=1.0):
    """A random solenoidal, Hermitian, dealiased spectral field triple in the
    reference layout, built from a random physical field through the oracle."""
    from oracle import oracle as O
    rng = np.random.default_rng(seed)
    L = N + 2
    kmax = kmax if kmax is not None else N // 4
    x = 2 * np.pi * np.arange(N) / N
    out = []
    comps = []
    for c in range(3):
        f = np.zeros((N, N, N))
        # a few random low modes: band-limited physical field
        for _ in range(12):
            k = rng.integers(-kmax, kmax + 1, size=3)
            ph = rng.uniform(0, 2 * np.pi)
            a = rng.standard_normal() * amp
            X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
            f += a * np.cos(k[0] * X + k[1] * Y + k[2] * Z + ph)
        phys = np.zeros((L, L, N), order="F")
        phys[:N, :N, :] = f
        comps.append(O.fft_r2c(phys))
    # project to solenoidal in spectral space (reference layout)
    i = np.arange(L)
    kz = np.where(i < N, np.where(i <= N // 2, i, i - N), 0).astype(float)
    kx = (np.arange(L) // 2).astype(float)
    k3 = np.arange(N)
    ky = np.where(k3 <= N // 2, k3, k3 - N).astype(float)
    KZ, KX, KY = np.meshgrid(kz, kx, ky, indexing="ij")
    KX[N:, :, :] = 0; KY[N:, :, :] = 0
    K2 = KX ** 2 + KY ** 2 + KZ ** 2
    K2[K2 == 0] = 1.0
    t = (KX * comps[0] + KY * comps[1] + KZ * comps[2]) / K2
    out = [np.asfortranarray(comps[0] - KX * t), np.asfortranarray(comps[1] - KY * t), np.asfortranarray(comps[2] - KZ * t)]
    for a in out:
        a[:, N:, :] = 0.0          # kx = N/2 plane
        a[N:, :, :] = 0.0          # padding
    return out


def rel_l2(a, b):
    num = sum(float(np.sum((x - y) ** 2)) for x, y in zip(a, b))
    den = sum(float(np.sum(y ** 2)) for y in b)
    return np.sqrt(num / max(den, 1e-300))


def spectral_mask_kx_lt_half(N):
    """Boolean mask of entries the GPU stores (kx < N/2, no padding rows)."""
    L = N + 2
    m = np.zeros((L, L, N), dtype=bool, order="F")
    m[:N, :N, :] = True
    return m


def random_droplets(n, seed, radius, domain=2 * np.pi):
    rng = np.random.default_rng(seed)
    yp = np.zeros((8, n), order="F")
    yp[0:3, :] = rng.uniform(0, domain, size=(3, n))
    r = np.asarray(radius, dtype=float)
    yp[6, :] = r if r.ndim == 0 else r[rng.integers(0, len(r), size=n)]
    yp[7, :] = np.arange(1, n + 1)
    return yp


def kgrid(N):
    """Wavenumbers of every entry of the reference spectral array (N+2, N+2, N):
    row i -> kz, column j -> kx = j // 2 (re/im pair), plane k -> ky.  Padding rows
    i >= N carry k = 0."""
    L = N + 2
    i = np.arange(L)
    kz = np.where(i < N, np.where(i <= N // 2, i, i - N), 0).astype(float)
    kx = (np.arange(L) // 2).astype(float)
    k3 = np.arange(N)
    ky = np.where(k3 <= N // 2, k3, k3 - N).astype(float)
    KZ, KX, KY = np.meshgrid(kz, kx, ky, indexing="ij")
    KX = KX.copy(); KY = KY.copy()
    KX[N:, :, :] = 0
    KY[N:, :, :] = 0
    return KZ, KX, KY


def to_spectral(oracle_mod, f, N):
    """Physical field f[x, y, z] (N^3) -> reference spectral array via the oracle FFT."""
    phys = np.zeros((N + 2, N + 2, N), order="F")
    phys[:N, :N, :] = f
    return oracle_mod.fft_r2c(phys)


def to_physical(oracle_mod, c, N):
    return oracle_mod.fft_c2r(c)[:N, :N, :]
