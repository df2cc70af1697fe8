"""The committed golden fixtures are consistent with the oracle that made
them: a fresh, independent low-spp FP64 oracle run of the Cornell box lies
within 3 sigma of tests/golden/cornell_fp64_4096spp.npz on almost every pixel."""
from pathlib import Path

import numpy as np

from ohao_engine_amd import scenes as S

GOLDEN = Path(__file__).with_name("golden")


def test_cornell_fixture_matches_fresh_oracle(oracle):
    ref = np.load(GOLDEN / "cornell_fp64_4096spp.npz")
    assert ref["mean"].shape == (256, 256, 3) and int(ref["spp"]) == 4096
    sc = S.cornell_box()
    o = oracle.OracleScene(sc)
    spp = 16
    rows = (96, 160)
    s, q = o.render(spp, 50_000, precision=64, rows=rows)
    # 8x8-pixel blocks: 1024 samples per block keep the z-scores Gaussian
    def blocks(a):
        return a.reshape(8, 8, 32, 8, 3).sum(axis=(1, 3))
    n = 64 * spp
    sb, qb = blocks(s[rows[0]:rows[1]]), blocks(q[rows[0]:rows[1]])
    m = sb / n
    v = np.maximum(qb / n - m ** 2, 0) * n / (n - 1)
    mr = ref["mean"][rows[0]:rows[1]].astype(np.float64).reshape(8, 8, 32, 8, 3).mean(axis=(1, 3))
    # variance of the block mean of the reference: per-pixel variances / (64 * 4096)
    vr = ref["var"][rows[0]:rows[1]].astype(np.float64).reshape(8, 8, 32, 8, 3).sum(axis=(1, 3)) / 64 ** 2 / 4096
    se = np.sqrt(v / n + vr)
    z = np.abs(m - mr) / se
    assert np.mean(z > 3.0) < 0.02
    assert np.max(z) < 6.0
    assert np.all(np.isfinite(ref["mean"])) and np.all(ref["var"] >= 0)

