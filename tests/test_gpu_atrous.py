"""ohpt_denoise_atrous (variance-guided a-trous, the reference's
DenoiseMode::Atrous comparator) against the numpy oracle on the same
accumulation buffers, on a size that is not a multiple of the 16x16 tile:
  * 16 spp (moment variance), 2 and 1 spp (spatial-variance bootstrap); at
    1 spp the display sums equal the display of the mean;
  * the call refuses a context that records no first-hit AOVs;
  * on a held-out view at 4 spp it lowers the display-space error of the
    plain accumulation against a 1024-spp render.
Tolerance: |GPU - oracle| <= 2e-4 per display value (FP32 vs float64 with
exp/pow of FP32 rounding), mean below 1e-5.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

from ohao_engine_amd import denoise as D
from ohao_engine_amd import ohpt, scenes as S

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import atrous_ref  # noqa: E402

pytestmark = pytest.mark.gpu


def small_scene(w, h, seed=11):
    return S.procedural(width=w, height=h, sphere_freq=8, hf_cells=(64, 32), n_instances=16, seed=seed,
                        env_size=(256, 128))


def sums_from_context(ctx):
    """Accumulation sums the filter reads, rebuilt from the readbacks (the
    sample count is a power of two, so the rescaling is exact)."""
    s, _ = ctx.read_moments()
    disp = ctx.read_display_sums()
    alb, nrm, dep = ctx.read_aovs()
    cnt = s[..., 3:4].astype(np.float64)
    albedo = np.concatenate([alb[..., :3] * cnt, dep[..., None] * cnt], -1)
    normal = nrm.astype(np.float64) * cnt
    return s, disp, albedo, normal


@pytest.mark.parametrize("spp", [16, 2, 1])
def test_atrous_matches_oracle(spp):
    W, H = 50, 37
    print(f"atrous parity {W}x{H} at {spp} spp", flush=True)
    sc = small_scene(W, H)
    ctx = ohpt.Context(W, H)
    try:
        ctx.load_scene(sc)
        D.with_guides(ctx, sc.settings)
        ctx.set_render_seed(5)
        ctx.reset()
        ctx.render(spp)
        out = ctx.denoise_atrous()
        acc, disp, alb, nrm = sums_from_context(ctx)
        mean = ctx.read_accum()
    finally:
        ctx.close()
    # the display sums are the per-sample display values the accumulation saw:
    # at one sample they equal the display of the mean
    if spp == 1:
        np.testing.assert_allclose(disp[..., :3], atrous_ref.display(mean[..., :3]), atol=2e-6)
    ref = atrous_ref.denoise_atrous(acc, disp, alb, nrm)
    assert np.all(np.isfinite(out))
    assert np.all(out[..., 3] == 1.0)
    err = np.abs(out[..., :3] - ref[..., :3])
    print(f"max |diff| {err.max():.2e}, mean {err.mean():.2e}", flush=True)
    assert err.max() <= 2e-4 and err.mean() < 1e-5
    # the filter did something: the output is not the unfiltered display image
    col, _, _, _ = atrous_ref.prepare(acc, disp, alb, nrm)
    assert np.abs(out[..., :3] - col).mean() > 1e-3


def test_atrous_needs_aovs():
    sc = small_scene(32, 32)
    ctx = ohpt.Context(32, 32)
    try:
        ctx.load_scene(sc)
        ctx.set_settings(sc.settings)   # flags without OHPT_FLAG_AOVS
        ctx.render(1)
        with pytest.raises(ohpt.OhptError, match="OHPT_FLAG_AOVS"):
            ctx.denoise_atrous()
    finally:
        ctx.close()


def test_atrous_lowers_display_error_of_a_short_accumulation():
    size = 128
    sc = S.procedural(width=size, height=size, seed=4242)
    ctx = ohpt.Context(size, size)
    try:
        ctx.load_scene(sc)
        D.with_guides(ctx, sc.settings)
        ctx.set_camera(D.random_view(np.random.default_rng(4242)))
        ctx.set_render_seed(7_000_000)
        ctx.reset()
        ctx.render(1024)
        ref = atrous_ref.display(ctx.read_accum()[..., :3])
        ctx.set_render_seed(3)
        ctx.reset()
        ctx.render(4)
        noisy = atrous_ref.display(ctx.read_accum()[..., :3])
        out = ctx.denoise_atrous()[..., :3]
    finally:
        ctx.close()
    e_noisy = float(np.sqrt(np.mean((noisy - ref) ** 2)))
    e_atr = float(np.sqrt(np.mean((out - ref) ** 2)))
    print(f"display RMSE at 4 spp: plain {e_noisy:.4f}, a-trous {e_atr:.4f}", flush=True)
    assert e_atr < 0.8 * e_noisy
