"""Scene generators of the measured workloads (ohao_engine_amd/scenes.py)."""
import numpy as np

from ohao_engine_amd import scenes as S


def test_heavy_scene_layout():
    """The heavy case's generator: 20 * freq^2 triangles per displaced sphere
    plus a 2-triangle ground; the spheres use the three textured channels."""
    sc = S.heavy(width=16, height=16, freq=6, n_instances=9, tex_size=64)
    assert sc.num_triangles == 9 * 20 * 36 + 2
    assert [t.shape for t in sc.textures] == [(64, 64, 3), (64, 64, 1), (64, 64, 1)]
    m = sc.materials[0]
    assert (m["albedo_tex"], m["roughness_tex"], m["metallic_tex"]) == (0, 1, 2)
    assert np.all(sc.mat[:-2] == 0) and np.all(sc.mat[-2:] == 1)
    assert sc.settings.max_bounces == 8
