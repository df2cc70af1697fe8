"""Closed-form pins of the oracle integrator (and hence of the estimator the GPU
kernels implement):
  * tilted-plane furnace: a flat plane under a constant environment reflects
    exactly its directional albedo E(mu) = int f(wo, wi) cos(wi) d wi, computed
    here by independent quadrature of the BSDF model (Lambert: E = rho);
  * emissive enclosure: a closed box whose walls emit Le and reflect rho
    (Lambert) has radiance Le * sum_{k=0..D} rho^k along paths of at most D
    bounces, with light sampling + MIS against BSDF sampling;
  * FP32 and FP64 oracle runs with the same random numbers agree.
"""
import math

import numpy as np
import pytest

from ohao_engine_amd import scenes as S


def ggx_albedo_quadrature(mu, rough, metal, albedo, f0d, n_t=1200, n_p=600):
    """E(mu) for the Lambert+GGX model (height-correlated Smith, Schlick F)."""
    a = rough * rough
    wo = np.array([math.sqrt(1 - mu * mu), 0.0, mu])
    th = (np.arange(n_t) + 0.5) / n_t * (np.pi / 2)
    ph = (np.arange(n_p) + 0.5) / n_p * (2 * np.pi)
    T, P = np.meshgrid(th, ph, indexing="ij")
    wi = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1)
    ci = wi[..., 2]
    h = wi + wo
    h /= np.linalg.norm(h, axis=-1, keepdims=True)
    ch = h[..., 2]
    voh = np.clip(h @ wo, 0, 1)
    D = a * a / (np.pi * (ch * ch * (a * a - 1) + 1) ** 2)

    def lam(c):
        c2 = c * c
        return 0.5 * (-1 + np.sqrt(1 + a * a * np.maximum(0, 1 - c2) / c2))

    G2 = 1.0 / (1 + lam(mu) + lam(ci))
    s = (1 - voh) ** 5
    F0 = f0d + (albedo - f0d) * metal
    F90 = 1.0 if (f0d > 0 or metal > 0) else 0.0
    F = F0 + (F90 - F0) * s
    f = F * D * G2 / (4 * mu * ci) + (1 - F) * (1 - metal) * albedo / np.pi
    dw = np.sin(T) * (np.pi / 2 / n_t) * (2 * np.pi / n_p)
    return float(np.sum(f * ci * dw))


def tilted_plane(theta_deg, material, W=24, H=24):
    th = math.radians(theta_deg)
    # plane through the origin with normal (0, sin th, cos th), large and facing +z camera
    n = np.array([0.0, math.sin(th), math.cos(th)])
    u = np.array([1.0, 0.0, 0.0])
    v = np.cross(n, u)
    s = 50.0
    c = [(-u - v) * s, (u - v) * s, (u + v) * s, (-u + v) * s]
    pos = np.array([np.concatenate([c[0], c[1], c[2]]), np.concatenate([c[0], c[2], c[3]])], np.float32)
    # winding so that the geometric normal is +n
    tri_n = np.cross(pos[0, 3:6] - pos[0, 0:3], pos[0, 6:9] - pos[0, 0:3])
    if np.dot(tri_n, n) < 0:
        pos = pos.reshape(2, 3, 3)[:, ::-1, :].reshape(2, 9).copy()
    cam = S.Camera.look_at((0.0, 0.0, 400.0), (0.0, 0.0, 0.0), 1.0)
    return S.Scene(pos=pos, mat=np.zeros(2, np.uint32), materials=np.array([material], S.MATERIAL_DTYPE),
                   env=S.constant_env(1.0), camera=cam, width=W, height=H,
                   settings=S.Settings(max_bounces=2, rr_start=1000, seed=2))


@pytest.mark.parametrize("case", [
    dict(rough=1.0, metal=0.0, albedo=0.6, f0=0.0, theta=30.0),    # Lambert: E = rho
    dict(rough=0.5, metal=1.0, albedo=1.0, f0=0.04, theta=40.0),   # GGX conductor furnace
    dict(rough=0.3, metal=0.0, albedo=0.5, f0=0.04, theta=60.0),   # dielectric: spec + (1-F) diffuse
    dict(rough=0.8, metal=0.5, albedo=0.7, f0=0.04, theta=20.0),
])
@pytest.mark.parametrize("flags", [0, S.FLAG_NO_ENV_NEE])
def test_plane_furnace_matches_quadrature(oracle, case, flags):
    m = S.material((case["albedo"],) * 3, roughness=case["rough"], metallic=case["metal"],
                   specular_f0=case["f0"])
    sc = tilted_plane(case["theta"], m)
    st = S.Settings(max_bounces=2, rr_start=1000, flags=flags, seed=4)
    o = oracle.OracleScene(sc, st)
    spp = 256
    s, q = o.render(spp, 0, precision=64)
    n = sc.width * sc.height * spp
    mean = s[..., 0].sum() / n
    var = q[..., 0].sum() / n - mean ** 2
    want = ggx_albedo_quadrature(math.cos(math.radians(case["theta"])), case["rough"], case["metal"],
                                 case["albedo"], case["f0"])
    if case["rough"] == 1.0 and case["f0"] == 0.0:
        assert want == pytest.approx(case["albedo"], rel=1e-4)
    assert abs(mean - want) < 5.0 * math.sqrt(var / n) + 2e-4 * want, (mean, want)


@pytest.mark.parametrize("D", [0, 1, 3])
@pytest.mark.parametrize("flags", [0, S.FLAG_NO_LIGHT_NEE])
def test_enclosure_closed_form(oracle, D, flags):
    Le = np.array([1.0, 0.5, 0.25])
    rho = 0.5
    sc = S.enclosure(tuple(Le), rho, max_bounces=D, width=16, height=16)
    st = S.Settings(max_bounces=D, rr_start=1000, flags=flags, seed=8)
    o = oracle.OracleScene(sc, st)
    spp = 128
    s, q = o.render(spp, 0, precision=64)
    n = sc.width * sc.height * spp
    mean = s.reshape(-1, 3).sum(0) / n
    var = q.reshape(-1, 3).sum(0) / n - mean ** 2
    want = Le * sum(rho ** k for k in range(D + 1))
    assert np.all(np.abs(mean - want) <= 5.0 * np.sqrt(var / n) + 1e-6), (mean, want)


def _reverse_winding(sc, inward_normals):
    """The enclosure with every triangle's winding reversed (facing out);
    optionally with flat vertex normals pointing in."""
    import dataclasses
    p = sc.pos.reshape(-1, 3, 3).astype(np.float64)
    n_in = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    n_in /= np.linalg.norm(n_in, axis=1, keepdims=True)
    pos = sc.pos.reshape(-1, 3, 3)[:, [0, 2, 1], :].reshape(-1, 9).copy()   # vertices 1 and 2 swapped
    nrm = np.repeat(n_in[:, None, :], 3, axis=1).reshape(-1, 9).astype(np.float32) if inward_normals else None
    return dataclasses.replace(sc, pos=pos, nrm=nrm)


@pytest.mark.parametrize("flags", [0, S.FLAG_NO_LIGHT_NEE])
def test_vertex_normals_define_the_outside(oracle, flags):
    # Inward-facing walls given by winding, or by vertex normals against an
    # outward winding, are the same enclosure: the hit side, the emitting side
    # and the sampled emitters' orientation follow the normals. Without the
    # normals the outward winding shows the camera only back faces: black.
    sc = S.enclosure((1.0, 0.5, 0.25), 0.5, max_bounces=2, width=12, height=12)
    st = S.Settings(max_bounces=2, rr_start=1000, flags=flags, seed=8)
    a, _ = oracle.OracleScene(sc, st).render(16, 0, precision=64)
    b, _ = oracle.OracleScene(_reverse_winding(sc, True), st).render(16, 0, precision=64)
    c, _ = oracle.OracleScene(_reverse_winding(sc, False), st).render(16, 0, precision=64)
    assert a.min() > 0
    np.testing.assert_allclose(b, a, rtol=1e-6)
    assert c.max() == 0.0


def test_russian_roulette_is_unbiased(oracle):
    Le = (1.0, 1.0, 1.0)
    sc = S.enclosure(Le, 0.6, max_bounces=6, width=12, height=12)
    a = oracle.OracleScene(sc, S.Settings(max_bounces=6, rr_start=1000, seed=1))
    b = oracle.OracleScene(sc, S.Settings(max_bounces=6, rr_start=1, seed=1))
    spp = 256
    n = 12 * 12 * spp
    sa, qa = a.render(spp, 0, precision=64)
    sb, qb = b.render(spp, 0, precision=64)
    ma, mb = sa[..., 0].sum() / n, sb[..., 0].sum() / n
    va, vb = qa[..., 0].sum() / n - ma ** 2, qb[..., 0].sum() / n - mb ** 2
    want = sum(0.6 ** k for k in range(7))
    assert abs(ma - want) < 5 * math.sqrt(va / n) + 1e-9
    assert abs(mb - want) < 5 * math.sqrt(vb / n)


def test_fp32_and_fp64_agree_with_same_random_numbers(oracle):
    sc = S.cornell_box(width=32, height=32)
    o = oracle.OracleScene(sc)
    s64, _ = o.render(16, 1, precision=64)
    s32, _ = o.render(16, 1, precision=32)
    rel = np.abs(s64 - s32).sum() / np.abs(s64).sum()
    assert rel < 2e-3


def test_sobol_sampler_converges_to_same_mean(oracle):
    sc = S.cornell_box(width=16, height=16)
    a = oracle.OracleScene(sc, S.Settings(max_bounces=8, rr_start=3, sampler=S.SAMPLER_PCG32, seed=1))
    b = oracle.OracleScene(sc, S.Settings(max_bounces=8, rr_start=3, sampler=S.SAMPLER_SOBOL_OWEN, seed=1))
    spp = 256
    sa, qa = a.render(spp, 0, precision=64)
    sb, qb = b.render(spp, 0, precision=64)
    ma, mb = sa.sum() / (16 * 16 * spp), sb.sum() / (16 * 16 * spp)
    va = qa.sum() / (16 * 16 * spp * 3) - (sa.sum() / (16 * 16 * spp * 3)) ** 2
    se = math.sqrt(max(va, 1e-12) / (16 * 16 * spp)) * 3
    assert abs(ma - mb) < 6 * se + 1e-3 * ma
