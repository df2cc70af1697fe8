"""Inverse fit on the path-replay adjoint (reduced size of the lantern
texture-fit workload): the adjoint texel gradient of the multi-view L2 loss
equals central finite differences of the same estimator (same samples) on
random visible texels, also at the reference's eps rule (2 % of the span),
and Adam steps on adjoint gradients lower the loss."""
import numpy as np
import pytest

from ohao_engine_amd import inverse as I

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fit():
    f = I.InverseFit(size=128, spp=8, views=2, tex_size=256, freq=30, target_spp=64)
    yield f
    f.close()


def test_adjoint_matches_central_fd(fit):
    res = I.run_check(fit, n_texels=16, eps=1e-3, seed=3)
    for r in res["rows"][:6]:
        print(f"  texel {r['texel']}: adjoint {r['adjoint']:.6e} fd {r['fd']:.6e} fd(2%) {r['fd_ref_eps']:.6e}",
              flush=True)
    print(f"median rel {res['median_rel_err']:.2e} max {res['max_rel_err']:.2e} cosine {res['cosine']:.6f}; "
          f"2% eps: median {res['median_rel_err_ref_eps']:.2e} max {res['max_rel_err_ref_eps']:.2e}", flush=True)
    assert res["median_rel_err"] < 1e-2 and res["max_rel_err"] < 5e-2
    assert res["cosine"] > 0.999
    assert res["median_rel_err_ref_eps"] < 3e-2


def test_adam_on_adjoint_recovers_texture(fit):
    hist, gpu_ms, observed = I.run_fit(fit, iters=12, lr=0.02)
    losses = [h["loss"] for h in hist if h["iter"] >= 0]
    errs = [h["texel_rmse_observed"] for h in hist]
    print("losses", [f"{x:.4f}" for x in losses], flush=True)
    print("observed-texel RMSE", [f"{x:.4f}" for x in errs], flush=True)
    assert losses[-1] < 0.5 * losses[0]
    assert errs[-1] < 0.7 * errs[0]


def test_heavy_case_workflow_small():
    """The combined heavy workflow (forward, adjoint and denoiser on every view,
    Adam on the device) at a reduced size: gradients reach all three textures
    and the parameters move."""
    h = I.HeavyCase(size=64, spp=2, views=2, tex=64, freq=6, instances=4, target_spp=8)
    try:
        th0 = [t.clone() for t in h.theta]
        r1 = h.iteration(lr=0.05)
        g = [float(t.abs().sum()) for t in h.grad]
        r2 = h.iteration(lr=0.05)
        print(r1, r2, g, flush=True)
        assert all(np.isfinite([r1["loss"], r2["loss"]])) and r1["loss"] > 0
        assert all(x > 0 for x in g)
        assert all(float((a - b).abs().max()) > 0 for a, b in zip(h.theta, th0))
        assert r1["forward_ms"] > 0 and r1["adjoint_ms"] > 0
    finally:
        h.close()
