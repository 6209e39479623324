"""Writes the FFT golden vectors used by tests/test_oracle_fft.py.

The expected outputs come from numpy.fft (pocketfft), an independent library DFT,
mapped into the reference's storage and normalisation:

  * physical field  v(x, y, z) in an (N+2, N+2, N) Fortran array, x, y < N used;
  * spectral field  v(kz, 2 kx + {0: re, 1: im}, ky) in the same array, kz rows
    0..N-1 (kz = i for i <= N/2, i - N above), kx = 0..N/2, ky columns 0..N-1;
  * forward transform  c(k) = N^-3 sum_x v(x) exp(-i k.x)  (mpifft.f90:81-142,
    1/N per stage), inverse unnormalised with exp(+i k.x) (mpifft.f90:147-198).

Run:  python tests/golden/make_fft_golden.py   (writes fft_golden.npz beside it)
"""
from pathlib import Path

import numpy as np


def to_reference_layout(c, N):
    """c[kx, ky, kz] (kx = 0..N/2) -> reference spectral array (N+2, N+2, N)."""
    L = N + 2
    out = np.zeros((L, L, N), order="F")
    for kx in range(N // 2 + 1):
        out[:N, 2 * kx, :] = c[kx, :, :].real.T
        out[:N, 2 * kx + 1, :] = c[kx, :, :].imag.T
    return out


def forward(phys, N):
    f = phys[:N, :N, :]                                  # f[x, y, z]
    c = np.fft.rfftn(f, axes=(1, 2, 0)) / N ** 3        # c[kx, ky, kz]
    return to_reference_layout(c, N)


def main():
    rng = np.random.default_rng(20240607)
    out = {}
    for N in (8, 16):
        L = N + 2
        phys = np.zeros((L, L, N), order="F")
        phys[:N, :N, :] = rng.standard_normal((N, N, N))
        out[f"phys_{N}"] = phys
        out[f"spec_{N}"] = forward(phys, N)
        # an inverse case: a random Hermitian spectrum from a second real field,
        # with the kx = N/2 plane cleared (the C2R reads it as zero)
        g = rng.standard_normal((N, N, N))
        c = np.fft.rfftn(g, axes=(1, 2, 0)) / N ** 3
        c[N // 2, :, :] = 0.0
        spec_in = to_reference_layout(c, N)
        phys_out = np.fft.irfftn(c, s=(N, N, N), axes=(1, 2, 0)) * N ** 3
        po = np.zeros((L, L, N), order="F")
        po[:N, :N, :] = phys_out
        out[f"inv_spec_{N}"] = spec_in
        out[f"inv_phys_{N}"] = po
    np.savez_compressed(Path(__file__).with_name("fft_golden.npz"), **out)


if __name__ == "__main__":
    main()
