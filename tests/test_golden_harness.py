"""The golden-image harness (tests/golden/render_golden.py) on the CPU:
PNG decoding against an independent statement of each PNG filter, the
reference's own golden as a real-world file, the resampling against closed
forms of Pillow's bilinear reduction, the pass rule, and the three modes
(check, --update, --selftest) driven by a stand-in render command."""
import json
import subprocess
import sys
import zlib
import struct
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLDEN))
import render_golden as G  # noqa: E402


def _encode_filtered(rgb, filters):
    """PNG spec encoder: each row filtered against the ORIGINAL bytes of the
    row above / to the left (filters[y] in 0..4)."""
    h, w, c = rgb.shape
    rows = rgb.reshape(h, w * c).astype(np.int32)
    out = b""
    for y in range(h):
        x = rows[y]
        up = rows[y - 1] if y else np.zeros_like(x)
        left = np.concatenate([np.zeros(c, np.int32), x[:-c]])
        ul = np.concatenate([np.zeros(c, np.int32), up[:-c]])
        f = filters[y]
        if f == 0:
            pred = np.zeros_like(x)
        elif f == 1:
            pred = left
        elif f == 2:
            pred = up
        elif f == 3:
            pred = (left + up) // 2
        else:
            p = left + up - ul
            pa, pb, pc = np.abs(p - left), np.abs(p - up), np.abs(p - ul)
            pred = np.where((pa <= pb) & (pa <= pc), left, np.where(pb <= pc, up, ul))
        out += bytes([f]) + ((x - pred) & 255).astype(np.uint8).tobytes()
    chunk = lambda k, b: struct.pack(">I", len(b)) + k + b + struct.pack(">I", zlib.crc32(k + b) & 0xFFFFFFFF)
    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0))
            + chunk(b"IDAT", zlib.compress(out)) + chunk(b"IEND", b""))


def test_png_decoder_against_every_filter(tmp_path):
    rng = np.random.default_rng(0)
    img = rng.integers(0, 256, (15, 11, 3), dtype=np.uint8)
    img[5:9] = np.clip(np.arange(11)[None, :, None] * 23 + np.arange(4)[:, None, None] * 7, 0, 255)  # smooth band
    filters = [y % 5 for y in range(15)]
    p = tmp_path / "f.png"
    p.write_bytes(_encode_filtered(img, filters))
    np.testing.assert_array_equal(G.read_png(p), img)
    G.write_png(tmp_path / "g.png", img)
    np.testing.assert_array_equal(G.read_png(tmp_path / "g.png"), img)


def test_reads_the_reference_golden():
    # the reference's committed 640x360 Cornell golden (PIL-encoded, Sub/Up/Paeth rows)
    img = G.read_png(GOLDEN / "reference_cornell_box.png")
    assert img.shape == (360, 640, 3)
    left, right = img[60:300, 20:60].reshape(-1, 3).mean(0), img[60:300, -60:-20].reshape(-1, 3).mean(0)
    assert left[0] > left[1] + 10 and right[1] > right[0]          # red wall left, green wall right
    # a mis-decoded filter row smears garbage down the image; this one is 16-spp noise
    assert np.abs(np.diff(img.astype(np.int32), axis=0)).mean() < 12


def test_bilinear_reduction_by_three_is_the_1_2_3_2_1_filter():
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, (9, 30, 3), dtype=np.uint8)
    out = G.resize_bilinear(img, 10, 9)
    k = np.array([1, 2, 3, 2, 1]) / 9.0
    ref = np.zeros((9, 10, 3))
    for o in range(10):
        xs = np.arange(3 * o - 1, 3 * o + 4)
        ok = (xs >= 0) & (xs < 30)
        ref[:, o] = np.tensordot(img[:, xs[ok]].astype(np.float64), k[ok] / k[ok].sum(), axes=([1], [0]))
    assert np.abs(out.astype(np.float64) - ref).max() <= 0.5 + 1e-6


def test_bilinear_reduction_sample_positions():
    # a unit ramp reduced 5x: output o samples the input at (o + 0.5) * 5 - 0.5
    ramp = np.broadcast_to(np.arange(250, dtype=np.uint8)[None, :, None], (2, 250, 3)).copy()
    out = G.resize_bilinear(ramp, 50, 2)
    np.testing.assert_array_equal(out[0, 1:-1, 0], 5 * np.arange(1, 49) + 2)
    const = np.full((1080, 1920, 3), 77, np.uint8)
    d = G.downscale(const)
    assert d.shape == (360, 640, 3) and np.all(d == 77)


def test_pass_rule():
    a = np.full((100, 100, 3), 100, np.int16)
    b = a.copy()
    b[0, 0, 1] += 4
    b[1:50, 0] -= 1                                   # 50 differing pixels = 0.5 %
    s = G.diff_stats(a, b)
    assert s["max_abs"] == 4 and s["diff_px"] == 50 and G.verdict(s, 4, 0.01)
    b[0, 0, 1] += 1
    assert not G.verdict(G.diff_stats(a, b), 4, 0.01)      # 5 LSB
    c = a.copy()
    c[:, :2] += 1                                     # 2 % of pixels by 1 LSB
    assert not G.verdict(G.diff_stats(a, c), 4, 0.01)
    assert G.diff_stats(a, a[:50]) is None


FAKE = ("import sys, numpy as np; sys.path.insert(0, {g!r}); import render_golden as G; "
        "rng = np.random.default_rng({seed}); img = np.full((720, 1280, 3), 120, np.uint8); "
        "img[100:200, 100:300] = rng.integers(0, 256, (100, 200, 3)); G.write_png(sys.argv[1], img)")


def _manifest(tmp_path, seed):
    cmd = f"{sys.executable} -c \"{FAKE.format(g=str(GOLDEN), seed=seed)}\" {{out}}"
    m = {"scenes": [{"name": "fake", "command": cmd, "golden": str(tmp_path / "fake.png")}]}
    path = tmp_path / f"manifest_{seed}.json"
    path.write_text(json.dumps(m))
    return path


def _run(manifest, *mode):
    r = subprocess.run([sys.executable, str(GOLDEN / "render_golden.py"), str(manifest), *mode],
                       capture_output=True, text=True, timeout=120)
    return r.returncode, r.stdout


def test_modes_end_to_end(tmp_path):
    m = _manifest(tmp_path, 1)
    rc, out = _run(m)
    assert rc == 1 and "golden missing" in out
    rc, out = _run(m, "--update")
    assert rc == 0 and G.read_png(tmp_path / "fake.png").shape == (360, 640, 3)
    rc, out = _run(m)
    assert rc == 0 and "[PASS]  fake: max_abs=0" in out
    rc, out = _run(m, "--selftest")
    assert rc == 0 and "(selftest)" in out
    rc, out = _run(_manifest(tmp_path, 2))            # different content under the same golden
    assert rc == 1 and "[FAIL]  fake" in out
    assert (tmp_path / "fake.actual.png").exists()
    m_bad = tmp_path / "bad.json"
    m_bad.write_text(json.dumps({"scenes": [{"name": "bad", "command": f"{sys.executable} -c \"raise SystemExit(3)\" {{out}}",
                                             "golden": str(tmp_path / "bad.png")}]}))
    rc, out = _run(m_bad)
    assert rc == 1 and "exit 3" in out
    rc, _ = _run(m, "--bogus")
    assert rc == 2
