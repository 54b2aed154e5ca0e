"""Seeded initial configurations for dense bidisperse suspensions.

Pure random sequential addition jams near phi ~ 0.38 in 3D, so dense states
(phi = 0.45 - 0.58) are made the way LF_DEM's own generator makes them
(GenerateInitConfig.cpp:37-154: random placement, then overlap relaxation):
uniform random placement at reduced radii, followed by radius inflation with
soft-sphere overlap relaxation until the largest overlap is below a tolerance.
Deterministic for a given seed.

Particle counts and box size follow GenerateInitConfig.cpp:548-572
(np1 from the volume-fraction ratio, rounded half-down; cubic box from the
actual particle volume).
"""
import math

import numpy as np


def bidisperse_counts(n, a1=1.0, a2=1.4, vf_ratio=0.5):
    pv1 = 4.0 / 3 * math.pi * a1 ** 3
    pv2 = 4.0 / 3 * math.pi * a2 ** 3
    vf1 = vf_ratio
    vf2 = 1 - vf_ratio
    total = n / (vf1 / pv1 + vf2 / pv2)
    t = vf1 * total / pv1
    n1 = int(t) if t - int(t) <= 0.5 else int(t) + 1
    return n1, n - n1


def box_size(radius, phi):
    pvol = float(np.sum(4.0 / 3 * math.pi * np.asarray(radius) ** 3))
    return (pvol / phi) ** (1.0 / 3)


def _pairs(pos, reach, L):
    """Pairs (i < j) closer than reach, sorted by (i, j)."""
    from scipy.spatial import cKDTree
    tree = cKDTree(np.mod(pos, L), boxsize=L)
    pr = tree.query_pairs(r=reach, output_type="ndarray")
    if len(pr):
        pr = np.sort(pr, axis=1)
        pr = pr[np.lexsort((pr[:, 1], pr[:, 0]))]
    return pr


def generate(n, phi, seed, a1=1.0, a2=1.4, overlap_tol=2e-3, max_iter=20000):
    """Returns (positions (n,3) in [0,L)^3, radii (n,), L) with max overlap
    (a_i + a_j - r_ij)/(a_i + a_j) below overlap_tol.

    The candidate pairs are a Verlet list: all pairs within the overlap
    reach plus a skin, rebuilt when the particles' largest accumulated
    displacement (twice, for two particles) plus the growth of the reach
    since the last build exceeds the skin, so every overlapping pair is in
    the list."""
    rng = np.random.default_rng(seed)
    n1, n2 = bidisperse_counts(n, a1, a2)
    radius = np.concatenate([np.full(n1, a1), np.full(n2, a2)])
    L = box_size(radius, phi)
    pos = rng.uniform(0, L, size=(n, 3))
    scale = 0.6
    it = 0
    step = 0.4
    skin = 0.6 * a2
    cand, moved, reach0 = None, 0.0, 0.0
    worst = 1.0
    while it < max_iter:
        it += 1
        r = radius * scale
        reach = 2 * r.max()
        if cand is None or 2 * moved + (reach - reach0) > skin:
            cand = _pairs(pos, reach + skin, L)
            moved, reach0 = 0.0, reach
        if len(cand) == 0:
            worst = 0.0
        else:
            i, j = cand[:, 0], cand[:, 1]
            d = pos[j] - pos[i]
            d -= L * np.round(d / L)
            dist = np.sqrt(np.einsum("ij,ij->i", d, d))
            contact = r[i] + r[j]
            ov = contact - dist
            m = ov > 0
            i, j, d, dist, ov, contact = i[m], j[m], d[m], dist[m], ov[m], contact[m]
            worst = float((ov / contact).max()) if len(ov) else 0.0
            if len(ov):
                u = d / np.maximum(dist, 1e-12)[:, None]
                push = (step * ov)[:, None] * u
                disp = np.empty_like(pos)
                for k in range(3):
                    disp[:, k] = np.bincount(j, push[:, k], minlength=n) - np.bincount(i, push[:, k], minlength=n)
                pos += disp
                pos %= L
                moved += float(np.sqrt(np.einsum("ij,ij->i", disp, disp).max()))
        if scale < 1.0:
            if worst < 0.05:
                scale = min(1.0, scale + 0.01)
        elif worst < overlap_tol:
            break
    if scale < 1.0 or worst >= overlap_tol:
        raise RuntimeError("confgen: relaxation did not converge (n=%d phi=%g seed=%d)" % (n, phi, seed))
    pos = np.mod(pos, L)
    pos[pos >= L] -= L
    return pos, radius, np.array([L, L, L])
