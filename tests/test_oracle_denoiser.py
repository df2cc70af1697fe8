"""Pins of the denoiser oracle (oracle/denoiser_ref.py): each building block
against a brute-force or closed-form statement, and the whole-network
gradient against central finite differences of its float64 forward."""
import sys
from pathlib import Path

import numpy as np
import torch

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import denoiser_ref as R  # noqa: E402


def test_conv3x3_layout_matches_brute_force():
    rng = np.random.default_rng(0)
    B, ci, co, H, W = 1, 3, 2, 5, 6
    x = rng.normal(size=(B, ci, H, W))
    w = rng.normal(size=(9, ci, co))
    b = rng.normal(size=co)
    y = R.conv3x3(torch.tensor(x), torch.tensor(w), torch.tensor(b), relu=False).numpy()
    ref = np.zeros((B, co, H, W))
    for o in range(co):
        for yy in range(H):
            for xx in range(W):
                s = b[o]
                for t in range(9):
                    dy, dx = t // 3 - 1, t % 3 - 1
                    sy, sx = yy + dy, xx + dx
                    if 0 <= sy < H and 0 <= sx < W:
                        s += np.dot(x[0, :, sy, sx], w[t, :, o])
                ref[0, o, yy, xx] = s
    assert np.allclose(y, ref, rtol=1e-12, atol=1e-12)


def test_upsample_half_pixel_weights():
    row = np.array([1.0, 5.0, 2.0])
    x = torch.tensor(np.tile(row, (2, 1))[None, None])
    y = R.upsample2(x).numpy()[0, 0, 0]
    a, b, c = row
    expect = [a, 0.75 * a + 0.25 * b, 0.25 * a + 0.75 * b, 0.75 * b + 0.25 * c, 0.25 * b + 0.75 * c, c]
    assert np.allclose(y, expect, rtol=0, atol=1e-12)


def test_kpn_zero_logits_is_replicated_box_mean():
    rng = np.random.default_rng(1)
    H, W = 7, 9
    rad = rng.uniform(0, 3, (1, 3, H, W))
    out = R.kpn_apply(torch.zeros(1, 25, H, W, dtype=torch.float64), torch.tensor(rad)).numpy()
    ref = np.zeros_like(rad)
    for y in range(H):
        for x in range(W):
            acc = np.zeros(3)
            for dy in range(-2, 3):
                for dx in range(-2, 3):
                    acc += rad[0, :, min(max(y + dy, 0), H - 1), min(max(x + dx, 0), W - 1)]
            ref[0, :, y, x] = acc / 25
    assert np.allclose(out, ref, rtol=1e-12, atol=1e-12)


def test_kpn_dominant_logit_selects_neighbour():
    rng = np.random.default_rng(2)
    H, W = 6, 8
    rad = torch.tensor(rng.uniform(0, 3, (1, 3, H, W)))
    k = 7   # dy = -1, dx = 0
    lg = torch.zeros(1, 25, H, W, dtype=torch.float64)
    lg[:, k] = 1000.0
    out = R.kpn_apply(lg, rad).numpy()
    y = np.arange(H)
    ref = rad.numpy()[:, :, np.clip(y - 1, 0, H - 1), :]
    assert np.allclose(out, ref, rtol=0, atol=1e-12)


def test_tonemap_known_points_and_derivative():
    d = lambda v: torch.tensor(v, dtype=torch.float64)
    assert abs(float(R.aces(d(1.0))) - 2.54 / 3.16) < 1e-12
    t0 = float(R.tonemap_loss(d(0.0)))
    assert abs(t0 - 1e-3 ** (1 / 2.2)) < 1e-12
    big = float(R.tonemap_loss(d(1e4)))
    assert abs(big - (1.0 + 1e-3) ** (1 / 2.2)) < 1e-12
    for x0 in (0.01, 0.3, 1.7, 5.0):
        x = torch.tensor(x0, dtype=torch.float64, requires_grad=True)
        R.tonemap_loss(x).backward()
        h = 1e-6
        fd = (float(R.tonemap_loss(d(x0 + h))) - float(R.tonemap_loss(d(x0 - h)))) / (2 * h)
        assert abs(float(x.grad) - fd) < 1e-7 * max(1.0, abs(fd))


def _guides(rng, B, H, W):
    rad = rng.lognormal(-1.0, 1.0, (B, H, W, 4))
    alb = rng.uniform(0.05, 0.95, (B, H, W, 4))
    nrm = rng.normal(size=(B, H, W, 4))
    nrm[..., :3] /= np.linalg.norm(nrm[..., :3], axis=-1, keepdims=True)
    dep = rng.uniform(0.5, 20.0, (B, H, W))
    tgt = rad * rng.uniform(0.7, 1.3, rad.shape)
    return rad, alb, nrm, dep, tgt


def test_network_gradient_matches_central_fd():
    rng = np.random.default_rng(3)
    rad, alb, nrm, dep, tgt = _guides(rng, 1, 16, 16)
    p = R.init_params(4).astype(np.float64)
    L, g = R.loss_and_grad(p, rad, alb, nrm, dep, tgt)
    sl = R.layer_slices()
    picks = []
    for li in (0, 5, 9, 10, 16, 18):
        ws, bs = sl[li]
        idx = np.arange(ws.start, ws.stop)
        big = idx[np.argsort(-np.abs(g[idx]))[:2]]
        picks += list(big) + [bs.start]
    eps = 1e-6
    for i in picks:
        q = p.copy()
        q[i] += eps
        lp, _ = R.loss_and_grad(q, rad, alb, nrm, dep, tgt)
        q[i] -= 2 * eps
        lm, _ = R.loss_and_grad(q, rad, alb, nrm, dep, tgt)
        fd = (lp - lm) / (2 * eps)
        assert abs(g[i] - fd) <= 1e-4 * abs(fd) + 1e-9, (i, g[i], fd)


def test_param_count_and_padding():
    assert R.param_count() == sum(9 * ci * co + co for ci, co in R.layer_shapes())
    assert R.layer_shapes()[10] == (160 + 128, 128)
    rng = np.random.default_rng(5)
    rad, alb, nrm, dep, _ = _guides(rng, 1, 10, 21)
    x = R.features(rad, alb, nrm, dep)
    assert x.shape == (1, 16, 16, 32)
    assert torch.equal(x[:, :, 9:, :21], x[:, :, 9:10, :21].expand(-1, -1, 7, -1))
    assert torch.all(x[:, 10:] == 0)
