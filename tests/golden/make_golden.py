"""Writes the golden vectors the oracle is pinned against (numpy / scipy,
computed here, independent of the oracle's code):

  fft_golden.npz     numpy.fft.fftn / ifftn*N of seeded complex arrays on small
                     grids covering radices 2, 3, 4, 5 and a prime length
  j1_golden.npz      scipy.special.j1 on the arguments the 2D soft-core kernel
                     (src/parallel.cu:411-437) evaluates on a 16x16 grid

Run:  python tests/golden/make_golden.py
"""
import os

import numpy as np
import scipy.special

HERE = os.path.dirname(os.path.abspath(__file__))
SHAPES = [(2, 3, 4), (4, 6, 8), (5, 3, 2), (8, 8, 8), (1, 7, 12), (3, 4, 9)]


def main():
    rng = np.random.default_rng(20240601)
    out = {}
    for i, s in enumerate(SHAPES):
        x = rng.standard_normal(s) + 1j * rng.standard_normal(s)
        out[f"x{i}"] = x
        out[f"f{i}"] = np.fft.fftn(x)
        out[f"b{i}"] = np.fft.ifftn(x) * x.size
    np.savez(os.path.join(HERE, "fft_golden.npz"), **out)

    # arguments k*a with k = 2 pi |k| on a 16x16 periodic grid of length 8, a = 1.5
    n, L, a = 16, 8.0, 1.5
    k1 = np.fft.fftfreq(n, d=L / n)
    KX, KY = np.meshgrid(k1, k1, indexing="ij")
    kk = 2 * np.pi * np.sqrt(KX ** 2 + KY ** 2)
    args = np.unique(np.round(kk[kk > 1e-6] * a, 15))
    np.savez(os.path.join(HERE, "j1_golden.npz"), args=args, j1=scipy.special.j1(args))


if __name__ == "__main__":
    main()
