"""ORACLE — TEST INFRASTRUCTURE ONLY.

CPU restatement (numpy, float64) of the variance-guided a-trous filter that
libohpt runs as ohpt_denoise_atrous: the SVGF spatial pass of the reference's
DenoiseMode::Atrous (ohao/render/rt/denoise/atrous_denoise.hpp:44-67,
shaders/rt/rt_svgf_atrous.comp:49-115, constants atrous_denoise.cpp:29-40).
Only tests/ use it, as the checker; nothing in the product imports it.

Inputs are the accumulation buffers of an offline render (per-pixel sums):
  accum   (H, W, 4)  .w = sample count (rgb unused)
  disp    (H, W, 4)  sum of the per-sample display value (rgb), sum of its
                     squared luminance (a)
  albedo  (H, W, 4)  sum of first-hit albedo, .w = sum of first-hit distance
  normal  (H, W, 4)  sum of first-hit normal, .w = number of first hits

Parity pins (tests/test_oracle_atrous.py):
  display map         the frame tonemap's ACES points
  moment variance     numpy var of the per-sample display luminances
  bootstrap variance  numpy var over 3x3 windows of an edge-padded image
  weights normalised  constant image and geometry pass through unchanged
  variance with w^2   constant variance shrinks by (sum b_i^2)^2 = (70/256)^2
                      per iteration, a closed form
  luminance stop off  huge variance on flat geometry = the plain B3 a-trous,
                      computed with scipy.ndimage.correlate (mode 'nearest')
  luminance stop on   zero variance keeps isolated luminance values
  depth stop          background (1e30) pixels never mix with geometry
  normal stop         opposite normals mix with weight exp(-2/0.0901)
"""
from __future__ import annotations

import numpy as np

LUMA = np.array([0.2126, 0.7152, 0.0722])
B3 = np.array([1.0, 4.0, 6.0, 4.0, 1.0]) / 16.0  # rt_svgf_atrous.comp:45
ITERATIONS = 5                                   # atrous_denoise.cpp:29
SIGMA_L, SIGMA_N, SIGMA_D = 0.4, 0.30, 2.0       # atrous_denoise.cpp:38-40
BACKGROUND_DEPTH = 1e30                          # rt_svgf_atrous.comp:100


def _aces(y):
    return y * (2.51 * y + 0.03) / (y * (2.43 * y + 0.59) + 0.14)


def display(x):
    """Display value of linear radiance: clamp(ACES(x / 2), 0, 1) ** (1/2.2)
    (the frame tonemap, pt_raygen_offline.rgen:102-105)."""
    a = np.clip(_aces(0.5 * np.asarray(x, np.float64)), 0.0, 1.0)
    return a ** (1.0 / 2.2)


def display_sums(samples):
    """(n, H, W, 3) linear radiance samples -> the (H, W, 4) display sums the
    renderer accumulates: sum of display(x) per channel, sum of luma^2."""
    t = display(samples)
    lum = t @ LUMA
    return np.concatenate([t.sum(0), (lum * lum).sum(0)[..., None]], -1)


def prepare(accum, disp, albedo, normal):
    """Filter inputs from the accumulation sums -> (colour (H,W,3), variance
    (H,W), normal (H,W,3), depth (H,W)), as the reference's temporal pass
    hands them over for a still camera (rt_svgf_temporal.comp:87-120):

    colour   = mean per-sample display value
    variance = E[l^2] - E[l]^2 of the per-sample display luminance l, raised
               to the 3x3 (edge-clamped) luminance variance of the colour
               where n < 4 (rt_svgf_temporal.comp:105-120)
    normal   = normalised mean first-hit normal
    depth    = mean first-hit distance over the hits, BACKGROUND_DEPTH when
               fewer than half the samples hit
    """
    cnt = np.asarray(accum, np.float64)[..., 3]
    inv = np.where(cnt > 0, 1.0 / np.where(cnt > 0, cnt, 1.0), 0.0)
    disp = np.asarray(disp, np.float64)
    col = disp[..., :3] * inv[..., None]
    m1 = col @ LUMA
    var = np.maximum(0.0, disp[..., 3] * inv - m1 * m1)
    H, W = var.shape
    yy, xx = np.arange(H), np.arange(W)
    s1 = np.zeros_like(var)
    s2 = np.zeros_like(var)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            t = _tap(m1, np.clip(yy + dy, 0, H - 1), np.clip(xx + dx, 0, W - 1))
            s1 += t
            s2 += t * t
    mm = s1 / 9.0
    var = np.where(cnt < 4, np.maximum(var, np.maximum(0.0, s2 / 9.0 - mm * mm)), var)
    nrm = np.asarray(normal, np.float64)
    nv = nrm[..., :3] * inv[..., None]
    ln = np.linalg.norm(nv, axis=-1, keepdims=True)
    nv = np.where(ln > 0, nv / np.where(ln > 0, ln, 1.0), nv)
    hits = nrm[..., 3]
    dsum = np.asarray(albedo, np.float64)[..., 3]
    ok = (2.0 * hits > cnt) & (hits > 0)
    depth = np.where(ok, dsum / np.where(hits > 0, hits, 1.0), BACKGROUND_DEPTH)
    return col, var, nv, depth


def _tap(a, ys, xs):
    return a[ys[:, None], xs[None, :]]


def atrous_iteration(col, var, nrm, depth, step, sigma_l=SIGMA_L, sigma_n=SIGMA_N, sigma_d=SIGMA_D):
    """One a-trous iteration, rt_svgf_atrous.comp:49-115 (clamp-to-edge taps)."""
    H, W = var.shape
    yy, xx = np.arange(H), np.arange(W)
    # 3x3 Gaussian prefilter of the variance (:62-72)
    gk = np.array([0.25, 0.5, 0.25])
    gvar = np.zeros_like(var)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            gvar += gk[dy + 1] * gk[dx + 1] * _tap(var, np.clip(yy + dy, 0, H - 1), np.clip(xx + dx, 0, W - 1))
    # (the shader divides by the weight sum, exactly 1 for these weights)
    sdev = np.sqrt(np.maximum(gvar, 1e-8))
    cl = col @ LUMA
    acc = np.zeros_like(col)
    wsum = np.zeros_like(var)
    vsum = np.zeros_like(var)
    for dy in range(-2, 3):
        for dx in range(-2, 3):
            ys = np.clip(yy + dy * step, 0, H - 1)
            xs = np.clip(xx + dx * step, 0, W - 1)
            sc = _tap(col, ys, xs)
            sn = _tap(nrm, ys, xs)
            sd = _tap(depth, ys, xs)
            sv = _tap(var, ys, xs)
            w = B3[dy + 2] * B3[dx + 2] * np.ones_like(var)
            w = w * np.exp(-np.abs(cl - sc @ LUMA) / (sdev * sigma_l + 1e-6))
            nd = np.maximum(1.0 - np.sum(nrm * sn, axis=-1), 0.0)
            w = w * np.exp(-nd / (sigma_n * sigma_n + 1e-4))
            w = w * np.exp(-np.abs(depth - sd) / (sigma_d * sigma_d + 1e-4))
            acc += sc * w[..., None]
            wsum += w
            vsum += w * w * sv
    ok = wsum > 1e-6
    safe = np.where(ok, wsum, 1.0)
    out_c = np.where(ok[..., None], acc / safe[..., None], col)
    out_v = np.where(ok, vsum / (safe * safe), gvar)
    return out_c, out_v


def atrous(col, var, nrm, depth, iterations=ITERATIONS, **sigmas):
    """Iterations with steps 1, 2, 4, ... -> (colour, variance)."""
    for it in range(iterations):
        col, var = atrous_iteration(col, var, nrm, depth, 1 << it, **sigmas)
    return col, var


def denoise_atrous(accum, disp, albedo, normal):
    """ohpt_denoise_atrous on the accumulation sums -> (H, W, 4), alpha 1."""
    col, var, nrm, depth = prepare(accum, disp, albedo, normal)
    out, _ = atrous(col, var, nrm, depth)
    return np.concatenate([out, np.ones(out.shape[:2] + (1,))], axis=-1)
