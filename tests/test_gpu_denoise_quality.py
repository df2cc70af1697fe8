"""What the trained denoiser gains, on a held-out procedural scene (seed 4242,
never used for training) at 1 spp, scored against a 2048-spp render:
  * trained network < untrained network (a near-box 5x5 filter) < raw 1 spp,
    in relMSE (eps 1e-2) and in LDR RMSE of the display transform;
  * the quality gate opens for the trained network and closes (falling back
    to plain accumulation) for a network whose relMSE exceeds the bound.
"""
import numpy as np
import pytest

from ohao_engine_amd import denoise as D
from ohao_engine_amd import ohpt, scenes as S

pytestmark = pytest.mark.gpu


def ldr(x):
    u = 0.5 * np.maximum(np.asarray(x, np.float64)[..., :3], 0.0)
    a = np.clip((u * (2.51 * u + 0.03)) / (u * (2.43 * u + 0.59) + 0.14), 0.0, 1.0)
    return a ** (1.0 / 2.2)


def ldr_rmse(x, ref):
    return float(np.sqrt(np.mean((ldr(x) - ldr(ref)) ** 2)))


@pytest.fixture(scope="module")
def heldout():
    if not D.WEIGHTS.exists():
        pytest.fail("trained weights missing from the package")
    size = 192
    sc = S.procedural(width=size, height=size, seed=4242)
    ctx = ohpt.Context(size, size)
    ctx.load_scene(sc)
    D.with_guides(ctx, sc.settings)
    ctx.set_camera(D.random_view(np.random.default_rng(4242)))
    ctx.set_render_seed(7_000_000)
    ctx.reset()
    ctx.render(2048)
    ref = ctx.read_accum()
    ctx.set_render_seed(3)
    ctx.reset()
    ctx.render(1)
    rad = ctx.read_accum()
    alb, nrm, dep = ctx.read_aovs()
    yield dict(ctx=ctx, ref=ref, rad=rad, alb=alb, nrm=nrm, dep=dep)
    ctx.close()


def test_trained_network_beats_untrained_and_raw(heldout):
    h = heldout
    trained = D.LearnedDenoiser(D.WEIGHTS)
    untrained = ohpt.Denoiser(seed=1)
    out_t = trained.dn.infer(h["rad"], h["alb"], h["nrm"], h["dep"])[0]
    out_u = untrained.infer(h["rad"], h["alb"], h["nrm"], h["dep"])[0]
    e_raw, e_u, e_t = D.relmse(h["rad"], h["ref"]), D.relmse(out_u, h["ref"]), D.relmse(out_t, h["ref"])
    l_raw, l_u, l_t = ldr_rmse(h["rad"], h["ref"]), ldr_rmse(out_u, h["ref"]), ldr_rmse(out_t, h["ref"])
    print(f"relMSE raw {e_raw:.4f} untrained {e_u:.4f} trained {e_t:.4f}; "
          f"LDR RMSE raw {l_raw:.4f} untrained {l_u:.4f} trained {l_t:.4f}", flush=True)
    # measured on this view: raw 0.2875, untrained 0.0293, trained 0.0139
    assert e_t < 0.75 * e_u and e_u < e_raw
    assert l_t < l_u < l_raw
    trained.close()
    untrained.close()


def test_quality_gate_opens_and_closes(heldout):
    h = heldout
    inputs = (h["rad"][None], h["alb"][None], h["nrm"][None], h["dep"][None])
    good = D.LearnedDenoiser(D.WEIGHTS)
    g = good.validate(inputs, h["ref"])
    print(f"trained: relMSE {g.relmse_denoised:.4f} (input {g.relmse_input:.4f}, bound {g.bound})", flush=True)
    assert g.enabled
    # a network that sharpens instead of filtering: large random head weights
    bad = D.LearnedDenoiser(weights=None)
    p = bad.dn.get_params()
    rng = np.random.default_rng(0)
    p = (p + rng.normal(0.0, 0.5, p.shape)).astype(np.float32)
    bad.dn.set_params(p)
    b = bad.validate(inputs, h["ref"])
    print(f"perturbed: relMSE {b.relmse_denoised:.4f} (bound {b.bound}) enabled {b.enabled}", flush=True)
    assert not b.enabled
    plain = bad.apply(h["ctx"])            # gate closed: plain accumulation
    assert np.array_equal(plain, h["ctx"].read_accum())
    good.close()
    bad.close()
