"""Writes the golden vectors that pin the FP64 oracle where closed forms do not.

The reference ships no tests or vectors (it is CUDA-only and cannot run here), so the fixtures
come from independent implementations available in this environment:
  fft_vectors.npz : numpy.fft forward/inverse transforms of seeded random vectors
  czt_vectors.npz : scipy.signal.czt on seeded random vectors, for the chirp-z parameters the
                    imaging chain uses (w = exp(-j 2 pi alpha / M), a = exp(-j pi alpha)) and
                    generic ones
Run from the repository root:  python tests/golden/make_golden.py
"""
from pathlib import Path

import numpy as np
import scipy
from scipy.signal import czt

HERE = Path(__file__).resolve().parent


def main() -> None:
    rng = np.random.default_rng(20260101)
    fft = {}
    for n in (8, 12, 48, 64, 192, 256, 384, 1000, 1024, 4096):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        fft[f"x_{n}"] = x
        fft[f"fwd_{n}"] = np.fft.fft(x)
        fft[f"inv_{n}"] = np.fft.ifft(x) * n          # unnormalised inverse
    np.savez_compressed(HERE / "fft_vectors.npz", **fft)

    cz = {}
    cases = [(64, 0.95, None, None), (128, 1.0, None, None), (192, 1.03, None, None), (256, 0.951, None, None),
             (50, None, np.exp(-1j * 0.3), np.exp(1j * 0.2)), (33, None, 1.01 * np.exp(-1j * 0.17), 0.98 + 0.1j)]
    for i, (m, alpha, w, a) in enumerate(cases):
        x = rng.standard_normal(m) + 1j * rng.standard_normal(m)
        if alpha is not None:
            w = np.exp(-2j * np.pi * alpha / m)
            a = np.exp(-1j * np.pi * alpha)
        cz[f"x_{i}"] = x
        cz[f"w_{i}"] = np.complex128(w)
        cz[f"a_{i}"] = np.complex128(a)
        cz[f"alpha_{i}"] = np.float64(np.nan if alpha is None else alpha)
        cz[f"out_{i}"] = czt(x, m=m, w=w, a=a)
    cz["scipy_version"] = np.array(scipy.__version__)
    np.savez_compressed(HERE / "czt_vectors.npz", **cz)
    print("wrote", HERE / "fft_vectors.npz", HERE / "czt_vectors.npz")


if __name__ == "__main__":
    main()
