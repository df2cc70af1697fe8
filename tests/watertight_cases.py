"""Ray/triangle cases on which a float edge function rounds to zero although
its exact value does not (the case the watertight test's exact-sign
re-evaluation exists for), with the hit decided by exact rational arithmetic.

Rays point straight down (direction (0, 0, -1)), so the sheared 2-D
coordinates of the watertight test are the vertex offsets themselves, exactly:
sheared x = world y, sheared y = world x. For an edge B-C through the ray's
foot point, U = Cx*By - Cy*Bx is built so that the two products round to the
same float while differing exactly; the third vertex A lies well off the edge,
so the other two edge functions have a definite sign and the ray hits exactly
when U has that sign too."""
from fractions import Fraction as F

import numpy as np

GRID = 2.0 ** -16


def _crafted_pairs(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        m = 1 << 16
        a = (m + rng.integers(1, m, 4096)) * GRID
        b = (m + rng.integers(1, m, 4096)) * GRID
        c = (m + rng.integers(1, m, 4096)) * GRID
        d = np.round(a * b / c / GRID) * GRID
        ok = (d >= 1.0) & (d < 2.0)
        for ai, bi, ci, di in zip(a[ok], b[ok], c[ok], d[ok]):
            fa, fb, fc, fd = (np.float32(v) for v in (ai, bi, ci, di))
            if fa * fb != fc * fd:              # float products differ: no rounding tie
                continue
            if F(float(fa)) * F(float(fb)) == F(float(fc)) * F(float(fd)):
                continue                         # exactly equal: not the case we want
            out.append((float(fa), float(fb), float(fc), float(fd)))
            if len(out) == n:
                break
    return out


def cases(n=128, seed=7):
    """-> list of (offsets (3, 3) float32 relative to the ray origin, expected_hit).

    Sheared coordinates: C = (a, c), B = (-d, -b), so U = Cx*By - Cy*Bx = c*d - a*b
    (float zero, exact nonzero); A lies off the line BC on alternating sides."""
    out = []
    for i, (a, b, c, d) in enumerate(_crafted_pairs(n, seed)):
        side = 1.0 if i % 2 == 0 else -1.0
        sA = (side * -2.0, side * 2.0)                 # sheared (x, y) of A
        sB = (-d, -b)
        sC = (a, c)
        # world offset (x, y, z) = (sheared y, sheared x, -1)
        off = np.array([[s[1], s[0], -1.0] for s in (sA, sB, sC)], np.float32)
        Ax, Ay = F(float(off[0, 1])), F(float(off[0, 0]))
        Bx, By = F(float(off[1, 1])), F(float(off[1, 0]))
        Cx, Cy = F(float(off[2, 1])), F(float(off[2, 0]))
        U = Cx * By - Cy * Bx
        V = Ax * Cy - Ay * Cx
        W = Bx * Ay - By * Ax
        assert U != 0 and abs(V) > F(1, 4) and abs(W) > F(1, 4)
        hit = (U > 0 and V > 0 and W > 0) or (U < 0 and V < 0 and W < 0)
        out.append((off, hit))
    return out


def scene_and_rays(n=128, seed=7):
    """The cases laid out on a grid (origins at multiples of 8, so vertex = origin + offset
    stays exact), one triangle per case, rays with tmax 2 so no ray reaches another case."""
    cs = cases(n, seed)
    pos = np.zeros((len(cs), 9), np.float32)
    rays = np.zeros((len(cs), 8), np.float32)
    hits = np.zeros(len(cs), bool)
    for i, (off, hit) in enumerate(cs):
        o = np.array([8.0 * (i % 8), 8.0 * ((i // 8) % 8), 4.0 * (i // 64)], np.float32)
        v = (o[None, :] + off).astype(np.float32)
        assert np.array_equal((v - o[None, :]).astype(np.float32), off)   # exact placement
        pos[i] = v.reshape(9)
        rays[i, 0:3] = o
        rays[i, 3] = 0.0
        rays[i, 4:7] = (0.0, 0.0, -1.0)
        rays[i, 7] = 2.0
        hits[i] = hit
    return pos, rays, hits
