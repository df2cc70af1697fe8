"""ORACLE — TEST INFRASTRUCTURE ONLY.

CPU restatement, in PyTorch on the CPU (float64), of the learned
kernel-predicting denoiser that libohpt runs on the GPU (ohpt_denoiser_*).
Only tests/, the smoke check and bench.py use it, as the checker; nothing in
the product imports it.

The reference has no learned denoiser (its denoise stage is Intel OIDN on the
CPU, ohao/render/rt/denoise/oidn_denoise.hpp:14-17, with the same
beauty/albedo/normal guides), so this module states the network the product
defines. Parity pins (tests/test_oracle_denoiser.py):
  conv3x3 weight layout   brute-force numpy correlation on a tiny input
  upsample x2             closed-form values of half-pixel bilinear weights
  max-pool + ReLU         torch.nn.functional (library routines)
  kpn filter              zero logits = 5x5 box mean with replicated edges;
                          a dominant logit selects the shifted neighbour
  loss tonemap            ACES fit at known points; derivative by central FD
  whole-network gradient  central finite differences of the float64 forward
With bf16=True the forward rounds features, weights and every stored
activation to bfloat16 at the points where the GPU stores them; gradients
through those casts are rounded to bfloat16 as well.
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F

# tensor table (name, level, channels) in evaluation order
TENSORS = [
    ("x0", 0, 16), ("a0", 0, 32), ("e0", 0, 32), ("p1", 1, 32), ("a1", 1, 64), ("e1", 1, 64),
    ("p2", 2, 64), ("a2", 2, 96), ("e2", 2, 96), ("p3", 3, 96), ("a3", 3, 128), ("e3", 3, 128),
    ("p4", 4, 128), ("a4", 4, 160), ("e4", 4, 160), ("u3", 3, 160), ("b3", 3, 128), ("d3", 3, 128),
    ("u2", 2, 128), ("b2", 2, 96), ("d2", 2, 96), ("u1", 1, 96), ("b1", 1, 64), ("d1", 1, 64),
    ("u0", 0, 64), ("b0", 0, 32), ("d0", 0, 32)]
# convolutions: (input, second input of a concatenation or None, output, cout, relu)
LAYERS = [
    ("x0", None, "a0", 32, True), ("a0", None, "e0", 32, True),
    ("p1", None, "a1", 64, True), ("a1", None, "e1", 64, True),
    ("p2", None, "a2", 96, True), ("a2", None, "e2", 96, True),
    ("p3", None, "a3", 128, True), ("a3", None, "e3", 128, True),
    ("p4", None, "a4", 160, True), ("a4", None, "e4", 160, True),
    ("u3", "e3", "b3", 128, True), ("b3", None, "d3", 128, True),
    ("u2", "e2", "b2", 96, True), ("b2", None, "d2", 96, True),
    ("u1", "e1", "b1", 64, True), ("b1", None, "d1", 64, True),
    ("u0", "e0", "b0", 32, True), ("b0", None, "d0", 32, True),
    ("d0", None, "logits", 25, False)]
CHAN = {n: c for n, _, c in TENSORS}
INPUT_CHANNELS = 10


def pad16(v: int) -> int:
    return (v + 15) & ~15


def layer_shapes():
    """(cin, cout) per convolution, cin counting both concatenated inputs."""
    out = []
    for a, b, _, cout, _ in LAYERS:
        out.append((CHAN[a] + (CHAN[b] if b else 0), cout))
    return out


def param_count() -> int:
    return sum(9 * ci * co + co for ci, co in layer_shapes())


def unpack(flat):
    """Flat parameter vector -> list of (W[9][cin][cout], b[cout]) tensors."""
    if not isinstance(flat, torch.Tensor):
        flat = torch.as_tensor(np.asarray(flat), dtype=torch.float64)
    res, off = [], 0
    for ci, co in layer_shapes():
        w = flat[off:off + 9 * ci * co].reshape(9, ci, co)
        off += 9 * ci * co
        b = flat[off:off + co]
        off += co
        res.append((w, b))
    assert off == flat.numel()
    return res


def init_params(seed: int, head_std: float = 0.3) -> np.ndarray:
    """Random parameters for tests (He-normal, small random bias, a head large
    enough that the predicted filters are far from a box)."""
    rng = np.random.default_rng(seed)
    parts = []
    for li, (ci, co) in enumerate(layer_shapes()):
        fan = 9 * (INPUT_CHANNELS if li == 0 else ci)
        std = np.sqrt(2.0 / fan) if li < len(LAYERS) - 1 else head_std * np.sqrt(1.0 / fan)
        w = rng.normal(0.0, std, (9, ci, co))
        if li == 0:
            w[:, INPUT_CHANNELS:, :] = 0.0
        parts += [w.ravel(), rng.normal(0.0, 0.05, co)]
    return np.concatenate(parts).astype(np.float32)


def bf16(x):
    return x.to(torch.bfloat16).to(torch.float64)


def _nchw(a):
    """(B, H, W, C) numpy/torch -> float64 (B, C, H, W)."""
    t = torch.as_tensor(np.asarray(a), dtype=torch.float64)
    if t.dim() == 3:
        t = t[None]
    return t.permute(0, 3, 1, 2).contiguous()


def features(rad, alb, nrm, dep, round_bf16=False):
    """Guides (B,H,W,4 / B,H,W) -> 16-channel input (B,16,Hp,Wp), edge-replicated."""
    r = _nchw(rad)[:, :3]
    B, _, H, W = r.shape
    a = _nchw(alb)[:, :3] if alb is not None else torch.zeros_like(r)
    n = _nchw(nrm)[:, :3] if nrm is not None else torch.zeros_like(r)
    if dep is not None:
        d = torch.as_tensor(np.asarray(dep), dtype=torch.float64).reshape(B, 1, H, W)
        d = torch.log1p(torch.clamp(d, min=0.0))
    else:
        d = torch.zeros((B, 1, H, W), dtype=torch.float64)
    x = torch.cat([torch.log1p(torch.clamp(r, min=0.0)), a, n, d, torch.zeros((B, 6, H, W), dtype=torch.float64)], 1)
    Hp, Wp = pad16(H), pad16(W)
    x = F.pad(x, (0, Wp - W, 0, Hp - H), mode="replicate")
    return bf16(x) if round_bf16 else x


def conv3x3(x, w, b, relu):
    """x (B,Cin,H,W); w[tap][ci][co] with tap = 3*dy + dx reading pixel
    (y+dy-1, x+dx-1); zero padding."""
    ci, co = w.shape[1], w.shape[2]
    weight = w.reshape(3, 3, ci, co).permute(3, 2, 0, 1)
    y = F.conv2d(x, weight, b, padding=1)
    return F.relu(y) if relu else y


def upsample2(x):
    return F.interpolate(x, scale_factor=2, mode="bilinear", align_corners=False)


def kpn_apply(logits, rad):
    """logits (B,25,H,W) over the 5x5 window (row-major, centre 12); rad (B,3,H,W)."""
    s = torch.softmax(logits, dim=1)
    B, _, H, W = rad.shape
    rp = F.pad(rad, (2, 2, 2, 2), mode="replicate")
    cols = F.unfold(rp, kernel_size=5).reshape(B, 3, 25, H, W)
    return (cols * s[:, None]).sum(2)


def aces(u):
    return (u * (2.51 * u + 0.03)) / (u * (2.43 * u + 0.59) + 0.14)


def tonemap_loss(x):
    """T(x) = (clamp(ACES(0.5 max(x,0)), 0, 1) + 1e-3)^(1/2.2)."""
    f = torch.clamp(aces(0.5 * torch.clamp(x, min=0.0)), 0.0, 1.0)
    return (f + 1e-3) ** (1.0 / 2.2)


def forward(params, rad, alb=None, nrm=None, dep=None, round_bf16=False, keep=False):
    """Returns (denoised (B,3,H,W), logits (B,25,Hp,Wp), activations dict)."""
    P = unpack(params) if not isinstance(params, list) else params
    rnd = bf16 if round_bf16 else (lambda t: t)
    r = _nchw(rad)[:, :3]
    H, W = r.shape[2], r.shape[3]
    act = {"x0": features(rad, alb, nrm, dep, round_bf16)}

    def conv(li):
        a, b2, o, _, relu = LAYERS[li]
        x = act[a] if b2 is None else torch.cat([act[a], act[b2]], 1)
        w, b = P[li]
        y = conv3x3(x, rnd(w), b, relu)
        act[o] = y if o == "logits" else rnd(y)

    for li in range(2):
        conv(li)
    act["p1"] = F.max_pool2d(act["e0"], 2)
    conv(2); conv(3)
    act["p2"] = F.max_pool2d(act["e1"], 2)
    conv(4); conv(5)
    act["p3"] = F.max_pool2d(act["e2"], 2)
    conv(6); conv(7)
    act["p4"] = F.max_pool2d(act["e3"], 2)
    conv(8); conv(9)
    act["u3"] = rnd(upsample2(act["e4"]))
    conv(10); conv(11)
    act["u2"] = rnd(upsample2(act["d3"]))
    conv(12); conv(13)
    act["u1"] = rnd(upsample2(act["d2"]))
    conv(14); conv(15)
    act["u0"] = rnd(upsample2(act["d1"]))
    conv(16); conv(17); conv(18)
    logits = act["logits"]
    out = kpn_apply(logits[:, :, :H, :W], r)
    if keep:
        for v in act.values():
            if v.requires_grad:
                v.retain_grad()
    return out, logits, act


def loss_fn(out, target):
    t = _nchw(target)[:, :3]
    return (tonemap_loss(out) - tonemap_loss(t)).abs().mean()


def loss_and_grad(params, rad, alb, nrm, dep, target, round_bf16=False):
    """Loss and its gradient w.r.t. the flat parameter vector (float64)."""
    p = torch.tensor(np.asarray(params, np.float64), requires_grad=True)
    out, logits, _ = forward(unpack(p), rad, alb, nrm, dep, round_bf16)
    L = loss_fn(out, target)
    L.backward()
    return float(L.detach()), p.grad.numpy().copy()


def layer_slices():
    """Slices of (W, b) of each convolution in the flat parameter vector."""
    res, off = [], 0
    for ci, co in layer_shapes():
        res.append((slice(off, off + 9 * ci * co), slice(off + 9 * ci * co, off + 9 * ci * co + co)))
        off += 9 * ci * co + co
    return res
