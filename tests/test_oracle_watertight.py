"""Watertightness of the oracle's FP32 ray/triangle test where a float edge
function rounds to zero: the decision must follow the exact sign (rational
arithmetic), and a straight-down ray through a shared vertex or edge of a
closed grid never slips through."""
import numpy as np

from watertight_cases import cases


def _tri_hit(oracle, ray, tri):
    import ctypes as C
    tuv = np.zeros(3, np.float32)
    p = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    r = np.ascontiguousarray(ray, np.float32)
    t = np.ascontiguousarray(tri, np.float32).reshape(9)
    return bool(oracle.lib().oracle_tri_hit(p(r), p(t), p(tuv))), tuv


def test_rounded_zero_edge_function_follows_exact_sign(oracle):
    cs = cases(128, seed=7)
    nhit = sum(h for _, h in cs)
    assert 0 < nhit < len(cs)            # both outcomes occur
    ray = np.array([0, 0, 0, 0, 0, 0, -1, np.inf], np.float32)
    for off, want in cs:
        got, tuv = _tri_hit(oracle, ray, off)
        assert got == want
        if got:
            assert abs(float(tuv[0]) - 1.0) <= 4e-7      # plane at distance 1, a few ulp


def test_grid_vertices_and_edges_never_leak(oracle):
    """A 6x6 grid of unit quads (two triangles each) at z = 0; rays straight down
    through every vertex, edge midpoint and face centre hit exactly once-or-more
    (never miss)."""
    tris = []
    for i in range(6):
        for j in range(6):
            p00, p10, p01, p11 = (i, j, 0), (i + 1, j, 0), (i, j + 1, 0), (i + 1, j + 1, 0)
            tris += [(p00, p10, p11), (p00, p11, p01)]
    tris = np.array(tris, np.float32)
    pts = [(x / 2, y / 2) for x in range(1, 12) for y in range(1, 12)]
    for (x, y) in pts:
        ray = np.array([x, y, 1, 0, 0, 0, -1, np.inf], np.float32)
        n = sum(_tri_hit(oracle, ray, t)[0] for t in tris)
        assert n >= 1, (x, y)
