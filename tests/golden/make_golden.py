"""Writes the committed golden vectors (numpy is the independent library):
small 3-D FFTs and a Poisson solve with the closed-form Green function.
Run: python tests/golden/make_golden.py"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(20261202)
    out = {}
    for G in (8, 16):
        z = rng.standard_normal((G, G, G)) + 1j * rng.standard_normal((G, G, G))
        out[f"fft_in_{G}"] = z
        out[f"fft_out_{G}"] = np.fft.fftn(z)
    # Poisson golden: G=16, L=40, r_s=2, deconvolution on, G const 1, rho = mass per cell
    G, L, rs = 16, 40.0, 2.0
    rho = rng.standard_normal((G, G, G))
    k = 2 * np.pi * np.fft.fftfreq(G, d=L / G)
    KZ, KY, KX = np.meshgrid(k, k, k, indexing="ij")
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    h = L / G
    W = (np.sinc(KX * h / 2 / np.pi) * np.sinc(KY * h / 2 / np.pi) * np.sinc(KZ * h / 2 / np.pi)) ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        green = -4 * np.pi / k2 * np.exp(-k2 * rs ** 2) / W ** 2 / h ** 3
    green[0, 0, 0] = 0
    out["poisson_rho"] = rho
    out["poisson_phi"] = np.real(np.fft.ifftn(np.fft.fftn(rho) * green))
    np.savez_compressed(os.path.join(HERE, "golden_fft_poisson.npz"), **out)


if __name__ == "__main__":
    main()
