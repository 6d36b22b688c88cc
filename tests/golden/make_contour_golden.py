"""Writes contour_golden.npz: phase-field slices and the contour radius
R(theta) that the reference's extract_contour_polar
(characteristicVelocity.py:75-122) computes from them.

The reference's algorithm runs on matplotlib's contour generator
(plt.contour, default algorithm).  oracle/postproc.contour_radius restates it
line by line -- all vertices of all paths, the centre at the mean of the
last path's vertices, the nearest polar angle per 3-degree step -- with
cs.allsegs[0] in place of the collections API that matplotlib 3.10 removed.
The fields cover one droplet (noisy, off-centre, elliptic), two droplets (the
last-path centre), a droplet cut by the array border (open lines), an
annulus (nested loops) and a smooth random field with saddles.
"""
import os
import sys

import matplotlib
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import postproc  # noqa: E402


def droplet(N, R0, xi, cx, cy, ax=1.0, ay=1.0):
    y, x = np.mgrid[0:N, 0:N].astype(np.float64)
    r = np.sqrt(((x - cx) / ax) ** 2 + ((y - cy) / ay) ** 2)
    return 0.5 * (1.0 - np.tanh((r - R0) / xi))


def smooth_random(rng, N, passes):
    f = rng.random((N, N))
    for _ in range(passes):
        f = (f + np.roll(f, 1, 0) + np.roll(f, 1, 1) + np.roll(f, (1, 1), (0, 1))) / 4
    return f


def fields():
    rng = np.random.default_rng(20261203)
    yield "noisy_droplet", droplet(64, 20.0, 2.0, 31.3, 32.7) + 0.02 * rng.standard_normal((64, 64))
    yield "ellipse", droplet(48, 1.0, 2.0 / 15.0, 22.6, 25.1, ax=15.0, ay=9.0)
    yield "two_droplets", droplet(48, 8.0, 1.5, 14.2, 13.7) + droplet(48, 10.0, 1.5, 32.4, 31.2)
    yield "border_droplet", droplet(48, 12.0, 2.0, 3.3, 25.8)
    yield "annulus", droplet(48, 17.0, 1.5, 24.3, 23.6) - droplet(48, 7.0, 1.5, 23.8, 24.4)
    yield "smooth_random", smooth_random(rng, 32, 3)
    yield "droplet_128", droplet(128, 40.0, 2.0, 64.2, 63.7) + 0.01 * rng.standard_normal((128, 128))


def main():
    out = {"matplotlib_version": np.array(matplotlib.__version__)}
    names = []
    for name, f in fields():
        f32 = f.astype(np.float32)
        R, c = postproc.contour_radius(f32.astype(np.float64), 0.5)
        out[f"field_{name}"] = f32
        out[f"R_{name}"] = R
        out[f"center_{name}"] = np.array(c)
        out[f"npaths_{name}"] = np.array(len(postproc.contour_paths(f32.astype(np.float64), 0.5)))
        names.append(name)
    out["names"] = np.array(names)
    np.savez_compressed(os.path.join(HERE, "contour_golden.npz"), **out)


if __name__ == "__main__":
    main()
