"""Writes dct_golden.npz: golden vectors for the DCT Poisson density force.

The reference's solver (neon_old_codebase/src/python/fluxion/force_electrostatic.py:
47-157) delegates its transforms to scipy.fft.dctn / idctn (type 2, norm
'ortho'; scipy is not pinned by the reference, scipy 1.15.3 here). This script
applies those library routines to the reference's published steps — point-in-
bin density, minus target, DCT, eigenvalue division, inverse DCT, centred
differences, nearest-bin gather — on seeded positions, and stores inputs and
outputs. It does not import the reference.
"""
import os

import numpy as np
from scipy.fft import dctn, idctn

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dct_golden.npz")


def solve(x, y, die_w, die_h, G, target=0.8, weight=1.0):
    bw, bh = die_w / G, die_h / G
    rho = np.zeros((G, G))
    cols = np.clip(x.astype(np.float64) / bw, 0, G - 1).astype(int)
    rows = np.clip(y.astype(np.float64) / bh, 0, G - 1).astype(int)
    for c, r in zip(cols, rows):
        rho[r, c] += 1.0 / (bw * bh)
    charge = rho - target
    wx, wy = 1.0 / bw, 1.0 / bh
    rh = dctn(charge, type=2, norm="ortho")
    kx, ky = np.meshgrid(np.arange(G), np.arange(G))
    ev = (wx * np.cos(np.pi * kx / G) - wx) ** 2 + (wy * np.cos(np.pi * ky / G) - wy) ** 2
    ev[0, 0] = 1.0
    ph = -rh / ev
    ph[0, 0] = 0.0
    phi = idctn(ph, type=2, norm="ortho")
    ex = np.zeros_like(phi)
    ey = np.zeros_like(phi)
    ex[:, 1:-1] = -(phi[:, 2:] - phi[:, :-2]) * wx * 0.5
    ey[1:-1, :] = -(phi[2:, :] - phi[:-2, :]) * wy * 0.5
    ex[:, 0] = -(phi[:, 1] - phi[:, 0]) * wx
    ex[:, -1] = -(phi[:, -1] - phi[:, -2]) * wx
    ey[0, :] = -(phi[1, :] - phi[0, :]) * wy
    ey[-1, :] = -(phi[-1, :] - phi[-2, :]) * wy
    fx = ex[rows, cols] * weight
    fy = ey[rows, cols] * weight
    return rho, phi, ex, ey, fx, fy


def main():
    rng = np.random.default_rng(20261203)
    data = {}
    for G, n, die in ((8, 60, 100.0), (16, 300, 250.0), (64, 2000, 1000.0)):
        x = rng.uniform(0, die, n).astype(np.float32)
        y = rng.uniform(0, die, n).astype(np.float32)
        x[:3] = [0.0, die, die * 0.5]  # edges and an interior bin boundary
        y[:3] = [die, 0.0, die * 0.5]
        rho, phi, ex, ey, fx, fy = solve(x, y, die, die, G)
        for k, v in dict(x=x, y=y, rho=rho, phi=phi, ex=ex, ey=ey, fx=fx, fy=fy).items():
            data[f"G{G}_{k}"] = v
        data[f"G{G}_die"] = np.array(die)
    np.savez_compressed(OUT, **data)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
