"""Pins of the oracle's samplers: reference KATs (tests/renderer/sobol_test.cpp),
bit-exact agreement with the reference C++ built into oracle/_ref, and the
published PCG32 demo vector."""
import ctypes as C

import numpy as np
import pytest


def test_pcg32_published_vector(oracle):
    # pcg32_srandom(42, 54): first outputs printed by O'Neill's pcg32-demo
    out = np.zeros(6, np.uint32)
    oracle.lib().oracle_pcg32_seq(42, 54, 6, out.ctypes.data_as(C.POINTER(C.c_uint32)))
    assert [hex(x) for x in out] == ["0xa15c02b7", "0x7b47f409", "0xba1d3330", "0x83d2f293", "0xbfa4784b",
                                     "0xcbed606e"]


# sobol_test.cpp:10-52
KAT_DIM01 = [(0.0, 0.0), (0.5, 0.5), (0.25, 0.75), (0.75, 0.25), (0.125, 0.625), (0.625, 0.125), (0.375, 0.375),
             (0.875, 0.875)]
KAT_DIM2 = [0.0, 0.5, 0.25, 0.75, 0.875, 0.375, 0.625, 0.125]
KAT_DIM3 = [0.0, 0.5, 0.75, 0.25, 0.125, 0.625, 0.875, 0.375]


def test_sobol_reference_kats(oracle):
    L = oracle.lib()
    for i in range(8):
        assert L.oracle_sobol_unscrambled(i, 0) == pytest.approx(KAT_DIM01[i][0], abs=1e-5)
        assert L.oracle_sobol_unscrambled(i, 1) == pytest.approx(KAT_DIM01[i][1], abs=1e-5)
        assert L.oracle_sobol_unscrambled(i, 2) == pytest.approx(KAT_DIM2[i], abs=1e-5)
        assert L.oracle_sobol_unscrambled(i, 3) == pytest.approx(KAT_DIM3[i], abs=1e-5)


def test_sobol_unit_interval_and_boundary(oracle):
    L = oracle.lib()
    for d in range(4):
        for i in range(128):
            v = L.oracle_sobol_unscrambled(i, d)
            assert 0.0 <= v < 1.0
    v = L.oracle_sobol_unscrambled(2 ** 31, 1)   # dirs[31] = 0xFFFFFFFF must stay < 1
    assert 0.0 <= v < 1.0


def test_owen_properties(oracle):
    L = oracle.lib()
    assert L.oracle_owen(0xABCD1234, 0xDEADBEEF) == L.oracle_owen(0xABCD1234, 0xDEADBEEF)
    assert L.oracle_owen(0x01234567, 1) != L.oracle_owen(0x01234567, 2)
    total = 0
    for i in range(1, 1001):
        v = (i * 0x9E3779B9) & 0xFFFFFFFF
        total += bin(L.oracle_owen(v, 1) ^ L.oracle_owen(v, 2)).count("1")
    assert 12000 < total < 20000


def test_owen_is_a_permutation_of_top_bits(oracle):
    """Owen scrambling is a nested uniform scramble: for fixed seed it maps the
    2^k values (i << (32-k)) bijectively onto 2^k distinct top-k-bit prefixes."""
    L = oracle.lib()
    k = 8
    tops = {L.oracle_owen(i << (32 - k), 0x1234567) >> (32 - k) for i in range(1 << k)}
    assert len(tops) == 1 << k


def test_sampler_matches_reference_cpp(oracle):
    R = oracle.ref()
    if R is None:
        pytest.skip("reference tree absent: oracle/_ref not built")
    L = oracle.lib()
    dirs = (C.c_uint32 * 128)()
    R.ref_sobol_dirs(dirs)
    for d in range(4):
        for i in list(range(512)) + [2 ** 31, 2 ** 31 + 77, 0xFFFFFFFF]:
            assert L.oracle_sobol_unscrambled(i, d) == R.ref_sobol_sample1d(i, d)
    rng = np.random.default_rng(1)
    for v, s in rng.integers(0, 2 ** 32, (4000, 2), dtype=np.uint64):
        assert L.oracle_owen(int(v), int(s)) == R.ref_owen(int(v), int(s))


def test_hash_pixel_known_values(oracle):
    """murmur3 fmix32 of x*c1 ^ y*c2 (sampler_sobol.glsl:51-58) recomputed in Python."""
    L = oracle.lib()
    M = 0xFFFFFFFF

    def h(x, y):
        v = ((x * 0x1b873593) & M) ^ ((y * 0xcc9e2d51) & M)
        v ^= v >> 16
        v = (v * 0x85ebca6b) & M
        v ^= v >> 13
        v = (v * 0xc2b2ae35) & M
        v ^= v >> 16
        return v

    for x, y in [(0, 0), (1, 0), (0, 1), (255, 17), (1919, 1079), (12345, 6789)]:
        assert L.oracle_hash_pixel(x, y) == h(x, y)


def test_sobol_owen_sample_layout(oracle):
    """getSample1D_sobol: pad = dim>>2 picks the Owen seed hashPixel ^ pad*0x9e3779b9."""
    L = oracle.lib()
    M = 0xFFFFFFFF
    for (px, py, idx, dim) in [(3, 5, 0, 0), (3, 5, 7, 2), (100, 20, 31, 9), (1, 2, 1000, 37)]:
        seed = L.oracle_hash_pixel(px, py) ^ (((dim >> 2) * 0x9e3779b9) & M)
        v = L.oracle_owen(L.oracle_sobol_int(idx, dim & 3), seed)
        assert L.oracle_sobol_sample(px, py, idx, dim) == np.float32((v >> 8) / 16777216.0)


def test_vertex_draws_are_stratified_sobol(oracle):
    """With the Sobol sampler, a 2D pair from one pad over 2^k consecutive sample
    indices is a (0,k,2)-net after Owen scrambling: one point per elementary box."""
    L = oracle.lib()
    out = np.zeros(9, np.float32)
    pts = []
    for s in range(64):
        # dims 2+0..8: dims (4, 5) are local dims (0, 1) of pad 1
        L.oracle_draw_vertex(1, 7, 9, 64, s, 0, out.ctypes.data_as(C.POINTER(C.c_float)))
        pts.append((out[2], out[3]))
    pts = np.array(pts)
    cells = set(zip((pts[:, 0] * 8).astype(int), (pts[:, 1] * 8).astype(int)))
    assert len(cells) == 64
