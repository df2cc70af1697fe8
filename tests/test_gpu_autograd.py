"""The adjoint through a torch.autograd.Function (tests/torch_render.py): on
the lantern scene at reduced size, loss.backward() gives the texel
gradient of a direct ohpt_adjoint call (deterministic mode, rtol 1e-5), and that gradient matches central
finite differences of the same estimator (same render seed) on visible
texels."""
import sys
from pathlib import Path

import numpy as np
import pytest
import torch

from ohao_engine_amd import ohpt, scenes as S

sys.path.insert(0, str(Path(__file__).resolve().parent))
from torch_render import render_texture  # noqa: E402

pytestmark = pytest.mark.gpu


def test_autograd_adjoint_matches_direct_call_and_fd():
    size, spp, seed = 64, 4, 77
    sc = S.lantern(width=size, height=size, spp=spp, seed=11, freq=20, tex_size=64)
    rc = ohpt.Context(size, size)
    try:
        rc.load_scene(sc)
        dev = torch.device("cuda:0")
        target = torch.from_numpy(sc.textures[0] * 0.5 + 0.25).to(dev)
        rc.update_texture(0, target)
        rc.set_render_seed(123)
        rc.reset()
        rc.render(64)
        ref = torch.empty((size, size, 4), device=dev)
        rc.read_accum_device(ref)
        tex = torch.full(sc.textures[0].shape, 0.5, device=dev, requires_grad=True)

        def loss_of(t):
            img = render_texture(t, rc, 0, spp, seed)
            return ((img[..., :3].double() - ref[..., :3].double()) ** 2).mean()

        loss = loss_of(tex)
        loss.backward()
        g_auto = tex.grad.detach().clone()
        # the same gradient by a direct call with dL/dI = 2 (img - ref) / (3 N)
        img = render_texture(tex.detach(), rc, 0, spp, seed)
        dl = torch.zeros_like(img)
        dl[..., :3] = (2.0 * (img[..., :3].double() - ref[..., :3].double()) / (3.0 * size * size)).float()
        torch.cuda.synchronize()
        rc.zero_gradients()
        rc.set_render_seed(seed)
        rc.adjoint(dl, spp, deterministic=True)
        g_direct = torch.empty_like(g_auto)
        rc.texture_gradient_device(0, g_direct)
        torch.cuda.synchronize()
        assert torch.allclose(g_auto, g_direct, rtol=1e-5, atol=1e-12)
        # central differences of the same estimator on the most influential texels
        flat = g_auto.reshape(-1).abs()
        picks = torch.topk(flat, 6).indices.tolist()
        eps = 1e-3
        for idx in picks:
            with torch.no_grad():
                tp, tm = tex.detach().clone().reshape(-1), tex.detach().clone().reshape(-1)
                tp[idx] += eps
                tm[idx] -= eps
                fd = (float(loss_of(tp.reshape(tex.shape))) - float(loss_of(tm.reshape(tex.shape)))) / (2 * eps)
            ga = float(g_auto.reshape(-1)[idx])
            print(f"texel {idx}: autograd {ga:.6e} fd {fd:.6e}", flush=True)
            assert abs(ga - fd) <= 2e-2 * abs(fd) + 1e-9
    finally:
        rc.close()
