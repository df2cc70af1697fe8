"""Pins of the oracle's detached finite differences (the reference derivative
the GPU path-replay adjoint is checked against): the FD is symmetric in eps,
converges as eps shrinks, and for a linear loss on a Lambert surface whose
only parameter is the albedo texel, d L / d albedo_c equals L_c / albedo_c on
the single-bounce image (radiance is linear in albedo there)."""
import numpy as np

from adjoint_scenes import oracle_detached_fd, textured_probe


def test_detached_fd_converges(oracle):
    sc = textured_probe(width=16, height=16, rough_metal=True)
    o = oracle.OracleScene(sc)
    w = np.random.default_rng(0).uniform(0.5, 1.5, (16, 16, 3))
    g1 = oracle_detached_fd(o, sc, w, 8, 0, (7, 3), 1, eps=1e-2)
    g2 = oracle_detached_fd(o, sc, w, 8, 0, (7, 3), 1, eps=1e-3)
    g3 = oracle_detached_fd(o, sc, w, 8, 0, (7, 3), 1, eps=1e-4)
    assert abs(g2 - g3) <= 1e-3 * max(abs(g3), 1e-6) + 1e-9
    assert abs(g1 - g3) <= 2e-2 * max(abs(g3), 1e-6) + 1e-9


def test_albedo_gradient_is_linear_on_single_bounce(oracle):
    """max_bounces=0 (primary hits only, no NEE reaches a second vertex):
    L is linear in the albedo texture, so the FD equals the finite quotient
    for any eps."""
    sc = textured_probe(width=16, height=16, rough_metal=False, max_bounces=1)
    o = oracle.OracleScene(sc)
    w = np.ones((16, 16, 3))
    ga = oracle_detached_fd(o, sc, w, 4, 0, (5, 9), 0, eps=1e-1)
    gb = oracle_detached_fd(o, sc, w, 4, 0, (5, 9), 0, eps=1e-3)
    assert abs(ga - gb) <= 1e-6 * max(abs(gb), 1e-9) + 1e-10
