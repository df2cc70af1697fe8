"""Test-harness wrapper of the path tracer as a differentiable torch op:
forward = ohpt_render of the context with the given texture (the mean image
as a device tensor), backward = ohpt_adjoint of the incoming image gradient
with the same render seed (the same samples), returning dL/d(texel) read on
the device. The product exposes only the C-ABI; this wrapper exists for the
tests."""
import torch


class RenderTexture(torch.autograd.Function):
    @staticmethod
    def forward(ctx, tex, rc, tid, spp, seed, deterministic=True):
        """tex: (h, w, c) float32 cuda tensor of texture `tid` of context `rc`;
        returns the (H, W, 4) mean image (alpha = sample count, no gradient)."""
        rc.update_texture(tid, tex.detach().contiguous())
        rc.set_render_seed(seed)
        rc.reset()
        rc.render(spp)
        img = torch.empty((rc.height, rc.width, 4), dtype=torch.float32, device=tex.device)
        rc.read_accum_device(img)
        ctx.rc, ctx.tid, ctx.spp, ctx.seed, ctx.det = rc, tid, spp, seed, deterministic
        ctx.tex_shape = tex.shape
        return img

    @staticmethod
    def backward(ctx, gout):
        rc = ctx.rc
        g = gout.detach().contiguous().float().clone()
        g[..., 3] = 0.0
        torch.cuda.synchronize()        # the library runs on its own stream
        rc.zero_gradients()
        rc.set_render_seed(ctx.seed)
        rc.adjoint(g, ctx.spp, deterministic=ctx.det)
        grad = torch.empty(ctx.tex_shape, dtype=torch.float32, device=gout.device)
        rc.texture_gradient_device(ctx.tid, grad)
        return grad, None, None, None, None, None


def render_texture(tex, rc, tid, spp, seed, deterministic=True):
    """Deterministic adjoint by default: bit-reproducible texel gradients."""
    return RenderTexture.apply(tex, rc, tid, spp, seed, deterministic)
