"""Writes tests/golden/golden.npz: library-routine vectors that pin the oracle.

* REDFT01 (FFTW's unnormalised DCT-III) == scipy.fft.dct(type=3, norm=None) (pocketfft),
  for the reference deck lengths 41, 81, 91, 121 and small/odd/even cases;
* unnormalised C2R == n * numpy.fft.irfftn on Hermitian half spectra (2D and 3D);
* K_nu(x) and Gamma(x) from scipy.special;
* the LCG known-answer test of SRC_F/normal_ScaRL.f90:762-769 (seed 12345).
Run: python tests/golden/gen_golden.py
"""
import os

import numpy as np
import scipy.fft as sf
import scipy.special as sp

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(20240601)
    out = {}
    for n in (1, 2, 3, 5, 8, 16, 41, 81, 91, 121, 144, 176):
        x = rng.standard_normal(n)
        out[f"dct3_in_{n}"] = x
        out[f"dct3_out_{n}"] = sf.dct(x, type=3, norm=None)
    for shp in ((8, 6, 1), (10, 8, 6), (16, 12, 9), (12, 7, 5)):
        n1, n2, n3 = shp
        # random complex half spectrum, made Hermitian by a real round trip
        f = rng.standard_normal((n3, n2, n1))
        H = np.fft.rfftn(f)  # C order (n3, n2, n1//2+1)
        key = "x".join(map(str, shp))
        out[f"c2r_in_{key}"] = np.transpose(H, (2, 1, 0)).copy()   # Fortran-ordered (h1, n2, n3) view
        out[f"c2r_out_{key}"] = np.transpose(np.fft.irfftn(H, s=(n3, n2, n1), axes=(0, 1, 2)) * (n1 * n2 * n3), (2, 1, 0)).copy()
    nus = np.array([0.0, 0.1, 0.3, 0.5, 1.0, 1.5, 2.7])
    xs = np.array([1e-3, 0.05, 0.5, 1.0, 3.0, 8.9, 9.1, 20.0, 40.0])
    out["kv_nu"] = nus
    out["kv_x"] = xs
    out["kv"] = np.array([[sp.kv(nu, x) for x in xs] for nu in nus])
    gx = np.array([0.1, 0.3, 0.5, 0.8, 1.3, 1.8, 2.5, 3.5, 7.3])
    out["gamma_x"] = gx
    out["gamma"] = sp.gamma(gx)
    out["lcg_seed"] = np.array([12345])
    out["lcg_seeds_out"] = np.array([207482415, 1790989824, 2035175616])
    out["lcg_r"] = np.array([0.096616, 0.833995, 0.947702])
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)


if __name__ == "__main__":
    main()
