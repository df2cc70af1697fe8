"""Oracle LBVH + traversal pins: Morton closed form, Karras invariants
(contiguous leaf ranges, parents' boxes contain children), the 4-wide
collapse (every leaf once, boxes nest), and the FP32 4-wide
traversal against an FP64 brute force over all triangles (closest primitive
equal except t-ties within 1e-5 relative, which are listed)."""
import numpy as np
import pytest

from ohao_engine_amd import scenes as S

LEAF = 0x80000000


def _interleave(q):
    r = 0
    for b in range(10):
        r |= ((q >> b) & 1) << (3 * b)
    return r


def test_morton_closed_form(oracle):
    import ctypes as C
    cmin = np.array([-1.0, 0.0, 2.0], np.float32)
    cmax = np.array([3.0, 8.0, 2.0], np.float32)   # zero z extent -> 0.5
    rng = np.random.default_rng(0)
    for _ in range(200):
        c = rng.uniform(cmin, cmax + [0, 0, 0]).astype(np.float32)
        got = oracle.lib().oracle_morton30(float(c[0]), float(c[1]), float(c[2]),
                                           cmin.ctypes.data_as(C.POINTER(C.c_float)),
                                           cmax.ctypes.data_as(C.POINTER(C.c_float)))
        q = []
        for a in range(3):
            ext = np.float32(cmax[a] - cmin[a])
            x = np.float32(0.5) if ext <= 0 else np.float32(np.float32(c[a] - cmin[a]) / ext)
            f = np.float32(x * np.float32(1024.0))
            q.append(0 if f <= 0 else (1023 if f >= 1023 else int(f)))
        want = (_interleave(q[0]) << 2) | (_interleave(q[1]) << 1) | _interleave(q[2])
        assert got == want


def random_soup(n, seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-5, 5, (n, 1, 3))
    v = c + rng.normal(0, 0.3 * scale, (n, 3, 3))
    return v.reshape(n, 9).astype(np.float32)


def _check_tree(nodes, tris, prim, keys, pos):
    n = pos.shape[0]
    ni = nodes.shape[0]
    boxes = nodes[:, :12].view(np.float32)
    child = nodes[:, 12:14]
    # keys sorted and unique, prim a permutation
    assert np.all(keys[1:] > keys[:-1])
    assert sorted(prim.tolist()) == list(range(n))
    # every leaf referenced exactly once (n > 1), every internal node except root once
    leaves = child[child >= LEAF] - LEAF
    inter = child[child < LEAF]
    if n > 1:
        assert sorted(leaves.tolist()) == list(range(n))
        assert sorted(inter.tolist()) == list(range(1, ni))
    # leaf ranges are contiguous and parent boxes contain child boxes
    rng_lo = np.zeros(ni, np.int64)
    rng_hi = np.zeros(ni, np.int64)

    def visit(i):
        lo, hi = [], []
        for k in range(2):
            c = int(child[i, k])
            b = boxes[i, 6 * k:6 * k + 6]
            if c >= LEAF:
                s = c - LEAF
                t = pos[prim[s]].reshape(3, 3)
                assert np.array_equal(b[:3], t.min(0)) and np.array_equal(b[3:], t.max(0))
                lo.append(s); hi.append(s)
            else:
                l, h = visit(c)
                cb = boxes[c]
                assert np.array_equal(b[:3], np.minimum(cb[0:3], cb[6:9]))
                assert np.array_equal(b[3:], np.maximum(cb[3:6], cb[9:12]))
                lo.append(l); hi.append(h)
        if n > 1:
            assert hi[0] + 1 == lo[1]
        return lo[0], hi[1]

    import sys
    sys.setrecursionlimit(10000)
    l, h = visit(0)
    assert (l, h) == (0, n - 1)


EMPTY = 0xFFFFFFFF


def _check_bvh4(nodes4, nodes, n):
    """4-wide collapse: one node per even-depth binary node, numbered in binary
    order; every leaf and every non-root node referenced once; slot boxes equal
    the binary child boxes, and an internal slot's box is the union of its
    node's slot boxes."""
    child = nodes[:, 12:14]
    boxes = nodes[:, :12].view(np.float32)
    ni = nodes.shape[0]
    depth = np.zeros(ni, np.int64)
    stack = [0]
    while stack:
        i = stack.pop()
        for c in child[i]:
            if c < LEAF:
                depth[c] = depth[i] + 1
                stack.append(int(c))
    even = np.flatnonzero(depth % 2 == 0)
    assert nodes4.shape[0] == even.size
    assert np.array_equal(nodes4[:, 28], even)                 # binary node of each 4-wide node, in order
    index_of = {int(b): q for q, b in enumerate(even)}
    lo = np.stack([nodes4[:, 0:4], nodes4[:, 4:8], nodes4[:, 8:12]], -1).view(np.float32)
    hi = np.stack([nodes4[:, 12:16], nodes4[:, 16:20], nodes4[:, 20:24]], -1).view(np.float32)
    refs = nodes4[:, 24:28]
    used = refs != EMPTY
    assert np.array_equal(nodes4[:, 29], used.sum(1))
    assert np.all(used[:, :2]) and np.all(used.sum(1) >= 2)
    assert np.all(used[:, :-1] | ~used[:, 1:])                # used slots first
    leaves = refs[used & (refs >= LEAF)] - LEAF
    inner = refs[used & (refs < LEAF)]
    if n > 1:
        assert sorted(leaves.tolist()) == list(range(n))
    assert sorted(inner.tolist()) == list(range(1, nodes4.shape[0]))
    for q, b in enumerate(even):
        want_box, want_ref = [], []
        for k in range(2):
            c = int(child[b, k])
            if c >= LEAF:
                want_box.append(boxes[b, 6 * k:6 * k + 6]); want_ref.append(c)
            else:
                for kk in range(2):
                    g = int(child[c, kk])
                    want_box.append(boxes[c, 6 * kk:6 * kk + 6])
                    want_ref.append(g if g >= LEAF else index_of[g])
        m = len(want_ref)
        assert refs[q, :m].tolist() == want_ref
        got = np.concatenate([lo[q, :m], hi[q, :m]], 1)
        assert np.array_equal(got.view(np.uint32), np.array(want_box).view(np.uint32))
        for k in range(m):
            r = int(refs[q, k])
            if r < LEAF:
                cu = used[r]
                assert np.array_equal(lo[q, k], lo[r][cu].min(0)) and np.array_equal(hi[q, k], hi[r][cu].max(0))


@pytest.mark.parametrize("n", [1, 2, 3, 17, 1000])
def test_lbvh_invariants(oracle, n):
    pos = random_soup(n, seed=n)
    sc = S.Scene(pos=pos, mat=np.zeros(n, np.uint32), materials=np.array([S.lambert((0.5,) * 3)]),
                 camera=S.Camera.look_at((0, 0, 20), (0, 0, 0), 40), width=4, height=4)
    o = oracle.OracleScene(sc)
    nodes, tris, prim, keys = o.bvh()
    _check_tree(nodes, tris, prim, keys, pos)
    _check_bvh4(o.bvh4(), nodes, n)


def test_lbvh_duplicate_centroids(oracle):
    """Many triangles with identical Morton codes: the index tiebreak keeps keys unique."""
    base = random_soup(1, seed=1)
    pos = np.repeat(base, 64, axis=0)
    sc = S.Scene(pos=pos, mat=np.zeros(64, np.uint32), materials=np.array([S.lambert((0.5,) * 3)]),
                 camera=S.Camera.look_at((0, 0, 20), (0, 0, 0), 40), width=4, height=4)
    o = oracle.OracleScene(sc)
    nodes, tris, prim, keys = o.bvh()
    _check_tree(nodes, tris, prim, keys, pos)
    _check_bvh4(o.bvh4(), nodes, 64)


def test_refit_keeps_topology_and_matches_brute_force(oracle):
    """Moving part of the geometry and refitting keeps the topology and leaf
    order, rebuilds every box exactly (the Karras invariants hold on the moved
    triangles) and traversal still finds the brute-force closest hit."""
    n = 1500
    pos = random_soup(n, seed=11)
    sc = S.Scene(pos=pos, mat=np.zeros(n, np.uint32), materials=np.array([S.lambert((0.5,) * 3)]),
                 camera=S.Camera.look_at((0, 0, 20), (0, 0, 0), 40), width=4, height=4)
    o = oracle.OracleScene(sc)
    nodes0, _, prim0, keys0 = o.bvh()
    moved = pos.copy().reshape(n, 3, 3)
    moved[n // 2:] = moved[n // 2:] * np.float32(2.0) + np.float32([1.5, -0.75, 0.25])
    moved = moved.reshape(n, 9)
    o.refit(moved)
    nodes1, tris1, prim1, keys1 = o.bvh()
    assert np.array_equal(nodes1[:, 12:14], nodes0[:, 12:14]) and np.array_equal(prim1, prim0)
    _check_tree(nodes1, tris1, prim1, keys1, moved)
    _check_bvh4(o.bvh4(), nodes1, n)
    rays = random_rays(3000, 23, origin_box=8.0)
    hits, _, _ = o.trace(rays)
    brute = o.brute(rays)
    bad, _ = compare_to_brute(moved, hits[:, 1].view(np.uint32), hits[:, 0], brute, rays)
    assert bad == []


def random_rays(n, seed, origin_box=6.0, tmin=0.0):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-origin_box, origin_box, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = tmin
    r[:, 4:7] = d
    r[:, 7] = np.inf
    return r


def _t64_single(pos, prim, ray):
    """FP64 distance along the ray to the plane of one triangle."""
    o = ray[0:3].astype(np.float64)
    d = ray[4:7].astype(np.float64)
    v = pos[prim].reshape(3, 3).astype(np.float64)
    e1, e2 = v[1] - v[0], v[2] - v[0]
    n = np.cross(e1, e2)
    den = n @ d
    if den == 0:
        return np.inf
    t = (n @ (v[0] - o)) / den
    return t


def compare_to_brute(pos, hits_prim, hits_t, brute, rays=None, rel_tie=1e-5):
    """Closest primitive vs the FP64 brute force. A different primitive is a
    t-tie (listed) when its own FP64 distance lies within rel_tie of the
    nearest one; anything else is a mismatch."""
    t64, p64, nties, ties = brute
    miss = 0xFFFFFFFF
    bad = []
    tie_list = []
    for i in range(len(p64)):
        if hits_prim[i] == p64[i]:
            if p64[i] != miss:
                assert abs(hits_t[i] - t64[i]) <= 1e-5 * max(1.0, abs(t64[i]))
            continue
        if p64[i] != miss and hits_prim[i] != miss:
            listed = set(ties[i, :min(nties[i], ties.shape[1])].tolist())
            if hits_prim[i] in listed:
                tie_list.append(i)
                continue
            if rays is not None and nties[i] > ties.shape[1]:
                tt = _t64_single(pos, int(hits_prim[i]), rays[i])
                if abs(tt - t64[i]) <= rel_tie * abs(t64[i]):
                    tie_list.append(i)
                    continue
        bad.append(i)
    return bad, tie_list


@pytest.mark.parametrize("scene_name", ["soup", "cornell"])
def test_trace_f32_vs_brute_f64(oracle, scene_name):
    if scene_name == "soup":
        pos = random_soup(400, seed=7)
        rays = random_rays(3000, 11)
    else:
        sc = S.cornell_box()
        pos = sc.pos
        rng = np.random.default_rng(5)
        o = np.tile(np.array([0.278, 0.273, -0.8]), (3000, 1)) + rng.uniform(-0.05, 0.05, (3000, 3))
        tgt = rng.uniform([0, 0, 0], [0.555, 0.555, 0.555], (3000, 3))
        d = tgt - o
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        rays = np.zeros((3000, 8), np.float32)
        rays[:, 0:3], rays[:, 4:7], rays[:, 7] = o, d, np.inf
    n = pos.shape[0]
    sc2 = S.Scene(pos=pos, mat=np.zeros(n, np.uint32), materials=np.array([S.lambert((0.5,) * 3)]),
                  camera=S.Camera.look_at((0, 0, 20), (0, 0, 0), 40), width=4, height=4)
    o = oracle.OracleScene(sc2)
    hits, visits, tests = o.trace(rays)
    prim = hits[:, 1].view(np.uint32)
    brute = o.brute(rays)
    bad, ties = compare_to_brute(pos, prim, hits[:, 0], brute, rays)
    assert bad == [], f"{len(bad)} rays disagree with the FP64 brute force"
    # occlusion agrees with closest-hit presence
    occ = o.occluded(rays)
    assert np.array_equal(occ.astype(bool), prim != 0xFFFFFFFF)
