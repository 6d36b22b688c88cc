"""Restatement of the reference's snapshot post-processing, for the tests.

TEST INFRASTRUCTURE ONLY: the checker of the library's diagnostics
(ttsh_get_centroid, ttsh_get_slice, ttsh_contour_radius, ttsh_run_ex's
diag.dat / a_R.txt).  Nothing in the product imports this file.

Fields here are in the library's device layout, [z][y][x].  The reference
reads a conf_<m>.dat snapshot into arrays of shape (Ly, Lx, Lz), i.e. axes
(y, x, z) (readDraw.py:205-215), so its first centroid coordinate is y.
"""
import numpy as np


def centroid(phi_zyx):
    """calculate_centroid (readDraw.py:30-52, characteristicVelocity.py:16-39):
    sum phi*i, phi*j, phi*k over the (y, x, z) axes over sum phi, and the
    truncations int().  Returns (idx, c, mass) with idx and c in (y, x, z)."""
    f = np.asarray(phi_zyx, np.float64)
    N = f.shape[0]
    r = np.arange(N, dtype=np.float64)
    m = f.sum()
    cz = (f.sum(axis=(1, 2)) * r).sum() / m
    cy = (f.sum(axis=(0, 2)) * r).sum() / m
    cx = (f.sum(axis=(0, 1)) * r).sum() / m
    c = (cy, cx, cz)
    return tuple(int(v) for v in c), c, m


def d1xO4(u, h):
    """op.py:3-7: 4th-order central d/dx along axis 1 (columns), periodic."""
    return (np.roll(u, 2, axis=1) - 8 * np.roll(u, 1, axis=1) + 8 * np.roll(u, -1, axis=1)
            - np.roll(u, -2, axis=1)) / (12 * h)


def d1yO4(u, h):
    """op.py:9-13: the same along axis 0 (rows)."""
    return (np.roll(u, 2, axis=0) - 8 * np.roll(u, 1, axis=0) + 8 * np.roll(u, -1, axis=0)
            - np.roll(u, -2, axis=0)) / (12 * h)


def slice_fields(phi, P, v, z, h):
    """Plane z as readDraw.py:220-237 takes it ([y][x] arrays) and the two
    vorticities d1xO4(vy) - d1yO4(vx), d1xO4(py) - d1yO4(px)."""
    g = lambda a: np.asarray(a[z], np.float64)
    out = {"phi": g(phi), "vx": g(v[0]), "vy": g(v[1]), "vz": g(v[2]),
           "px": g(P[0]), "py": g(P[1]), "pz": g(P[2])}
    out["vort_v"] = d1xO4(out["vy"], h) - d1yO4(out["vx"], h)
    out["vort_p"] = d1xO4(out["py"], h) - d1yO4(out["px"], h)
    return out


def contour_paths(field, level=0.5):
    """The vertex arrays of plt.contour(field, levels=[level]) in path order,
    as characteristicVelocity.py:77-79 obtains them.  The reference reads
    contour_set.collections[0].get_paths(), an API matplotlib 3.10 removed;
    cs.allsegs[0] holds the same vertex arrays.  Needs matplotlib (default
    contour algorithm)."""
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    cs = plt.contour(np.asarray(field, np.float64), levels=[level], colors="black")
    plt.close()
    return [np.asarray(s, np.float64) for s in cs.allsegs[0] if len(s)]


def contour_radius(field, level=0.5):
    """extract_contour_polar (characteristicVelocity.py:75-122): all vertices
    of all paths in order, the centre at the mean of the LAST path's vertices
    (:85, inside the loop, so the last iteration wins), and for theta = -180,
    -177, ..., 177 degrees the radius of the first vertex with the nearest
    polar angle (np.argmin).  Returns (R[120], (cx, cy))."""
    paths = contour_paths(field, level)
    if not paths:
        raise ValueError("the level is not crossed")
    pts = np.concatenate(paths)
    cx, cy = np.mean(paths[-1], axis=0)
    dx, dy = pts[:, 0] - cx, pts[:, 1] - cy
    pr = np.sqrt(dx ** 2 + dy ** 2)
    pt = np.arctan2(dy, dx)
    R = np.empty(120)
    for a, deg in enumerate(range(-180, 180, 3)):
        R[a] = pr[np.argmin(np.abs(pt - deg / 180 * np.pi))]
    return R, (float(cx), float(cy))
