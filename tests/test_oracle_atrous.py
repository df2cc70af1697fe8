"""Pins of the a-trous oracle (oracle/atrous_ref.py) against closed forms and
library routines, so that a dropped term, a wrong tap index or a wrong sign in
an edge stop fails one of them."""
import sys
from pathlib import Path

import numpy as np
from scipy import ndimage

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import atrous_ref as A  # noqa: E402


def _flat_geometry(H, W, depth=3.0):
    nrm = np.zeros((H, W, 3))
    nrm[..., 2] = 1.0
    return nrm, np.full((H, W), depth)


def test_display_map_known_points():
    # ACES(1) = 2.54 / 3.16 (x = 2 maps to y = 1), ACES(0) = 0, clamp at 1
    assert abs(A.display(2.0) - (2.54 / 3.16) ** (1 / 2.2)) < 1e-12
    assert A.display(0.0) == 0.0
    assert A.display(1e6) == 1.0


def test_prepare_colour_and_variance_are_the_per_sample_display_moments():
    rng = np.random.default_rng(1)
    n = 64
    samples = rng.gamma(0.7, 0.8, size=(n, 2, 3, 3))    # heavy-ish tail: some clamp at 1
    accum = np.concatenate([samples.sum(0), np.full((2, 3, 1), n)], -1)
    col, var, _, _ = A.prepare(accum, A.display_sums(samples), np.zeros((2, 3, 4)), np.zeros((2, 3, 4)))
    t = A.display(samples)
    np.testing.assert_allclose(col, t.mean(0), rtol=1e-12)
    np.testing.assert_allclose(var, (t @ A.LUMA).var(0), rtol=1e-9)   # population variance


def test_prepare_bootstraps_short_accumulations_with_the_spatial_variance():
    rng = np.random.default_rng(4)
    H, W = 5, 7
    for n in (1, 3, 4):
        samples = np.broadcast_to(rng.uniform(0, 2, (1, H, W, 3)), (n, H, W, 3))   # zero moment variance
        accum = np.concatenate([samples.sum(0), np.full((H, W, 1), n)], -1)
        col, var, _, _ = A.prepare(accum, A.display_sums(samples), np.zeros((H, W, 4)), np.zeros((H, W, 4)))
        lum = np.pad(col @ A.LUMA, 1, mode="edge")
        win = np.lib.stride_tricks.sliding_window_view(lum, (3, 3)).reshape(H, W, 9)
        ref = win.var(-1) if n < 4 else np.zeros((H, W))
        np.testing.assert_allclose(var, ref, rtol=1e-9, atol=1e-12)


def test_prepare_geometry_and_background_rule():
    accum = np.zeros((1, 3, 4))
    accum[..., 3] = 8
    nrm = np.zeros((1, 3, 4))
    alb = np.zeros((1, 3, 4))
    nrm[0, 0] = [0, 0, 16, 8]            # every sample hit, normal +z (x2 scale)
    alb[0, 0, 3] = 8 * 2.5               # mean hit distance 2.5
    nrm[0, 1, 3] = 4                     # exactly half hit: background
    nrm[0, 2] = [3, 0, 0, 5]             # 5 of 8 hit: depth over the hits
    alb[0, 2, 3] = 5 * 7.0
    _, _, n, d = A.prepare(accum, np.zeros((1, 3, 4)), alb, nrm)
    np.testing.assert_allclose(n[0, 0], [0, 0, 1])
    np.testing.assert_allclose(n[0, 2], [1, 0, 0])
    assert d[0, 0] == 2.5 and d[0, 1] == A.BACKGROUND_DEPTH and d[0, 2] == 7.0


def test_constant_image_passes_and_variance_shrinks_by_the_squared_weights():
    H, W = 9, 11
    col = np.broadcast_to(np.array([0.3, 0.5, 0.7]), (H, W, 3)).copy()
    var = np.full((H, W), 0.02)
    nrm, dep = _flat_geometry(H, W)
    c, v = A.atrous_iteration(col, var, nrm, dep, 2)
    np.testing.assert_allclose(c, col, rtol=1e-12)
    np.testing.assert_allclose(v, 0.02 * (70 / 256) ** 2, rtol=1e-12)
    c, v = A.atrous(col, var, nrm, dep)
    np.testing.assert_allclose(v, 0.02 * (70 / 256) ** 10, rtol=1e-10)


def _b3_atrous_scipy(img, iterations):
    out = img.copy()
    for it in range(iterations):
        s = 1 << it
        k1 = np.zeros(4 * s + 1)
        k1[::s] = A.B3
        k = np.outer(k1, k1)
        out = np.stack([ndimage.correlate(out[..., c], k, mode="nearest") for c in range(3)], -1)
    return out


def test_huge_variance_on_flat_geometry_is_the_plain_b3_atrous():
    rng = np.random.default_rng(2)
    H, W = 13, 17                        # smaller than the widest step: clamped taps matter
    col = rng.uniform(0, 1, (H, W, 3))
    var = np.full((H, W), 1e30)
    nrm, dep = _flat_geometry(H, W)
    c, _ = A.atrous(col, var, nrm, dep, iterations=5)
    np.testing.assert_allclose(c, _b3_atrous_scipy(col, 5), rtol=0, atol=1e-9)


def test_zero_variance_keeps_isolated_values():
    H, W = 8, 8
    col = np.full((H, W, 3), 0.4)
    col[3, 4] = [0.9, 0.9, 0.9]          # luminance differs by 0.5 from every neighbour
    nrm, dep = _flat_geometry(H, W)
    c, _ = A.atrous(col, np.zeros((H, W)), nrm, dep)
    np.testing.assert_allclose(c[3, 4], [0.9, 0.9, 0.9], atol=1e-12)
    np.testing.assert_allclose(c[0, 0], [0.4, 0.4, 0.4], atol=1e-12)


def test_depth_stop_separates_background_from_geometry():
    rng = np.random.default_rng(3)
    H, W = 12, 12
    nrm, dep = _flat_geometry(H, W)
    dep[:, 6:] = A.BACKGROUND_DEPTH
    col = rng.uniform(0, 1, (H, W, 3))
    var = np.full((H, W), 0.5)
    # one iteration: later ones see the filtered variance, whose 3x3 prefilter
    # (rt_svgf_atrous.comp:62-72) has no edge stop, so only this one is exact
    c1, _ = A.atrous_iteration(col, var, nrm, dep, 2)
    col2 = col.copy()
    col2[:, :6] = rng.uniform(0, 1, (H, 6, 3))   # change only the geometry side
    c2, _ = A.atrous_iteration(col2, var, nrm, dep, 2)
    assert np.array_equal(c1[:, 6:], c2[:, 6:])
    assert not np.allclose(c1[:, :6], c2[:, :6])


def test_normal_stop_weight_of_opposite_normals():
    # two pixels, opposite normals, equal depth, huge variance: the centre's
    # output = (w0 c0 + w1 c1) / (w0 + w1) with the tap weight of the other
    # pixel scaled by exp(-2 / (0.3^2 + 1e-4))
    H, W = 1, 2
    col = np.array([[[1.0, 1.0, 1.0], [0.0, 0.0, 0.0]]])
    nrm = np.array([[[0, 0, 1.0], [0, 0, -1.0]]])
    dep = np.full((H, W), 1.0)
    var = np.full((H, W), 1e30)
    c, _ = A.atrous_iteration(col, var, nrm, dep, 1)
    # taps clamp: for pixel 0, dx in {-2,-1,0} land on pixel 0, {1,2} on pixel 1
    own = A.B3[:3].sum()
    other = A.B3[3:].sum() * np.exp(-2.0 / (0.09 + 1e-4))
    np.testing.assert_allclose(c[0, 0, 0], own / (own + other), rtol=1e-9)
