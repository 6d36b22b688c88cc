"""Writes fft_golden.npz: seeded complex inputs and numpy's DFTs of them.

numpy.fft is a third-party implementation of the same unnormalised DFT that
cufftExecZ2Z computes (forward e^{-2 pi i jk/N}; inverse e^{+2 pi i jk/N},
numpy's ifftn times N^d).  The reference has no FFT test vectors of its own,
so these fixtures pin the oracle's FFT against an independent implementation
at radix-2 and radix-3 sizes.
"""
import os
import numpy as np

out = {}
rng = np.random.default_rng(20261202)
for N in (8, 12, 16, 48, 96):
    a = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    out[f"x1_{N}"] = a
    out[f"f1_{N}"] = np.fft.fft(a)
    out[f"i1_{N}"] = np.fft.ifft(a) * N
for N in (6, 8, 12):
    a = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    out[f"x3_{N}"] = a
    out[f"f3_{N}"] = np.fft.fftn(a)
    out[f"i3_{N}"] = np.fft.ifftn(a) * N ** 3
np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_golden.npz"), **out)
