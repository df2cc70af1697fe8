#!/usr/bin/env python3
"""Golden-image regression harness for the example drivers.

Same contract as the reference's harness (tests/golden/render_golden.py:14-23,
:33-35, :52-164): every scene of a manifest is rendered by its command (with
{out} replaced by a temporary PNG path), downscaled to 640 wide with Pillow's
BILINEAR resampling, and compared with the committed golden PNG. A scene
passes when BOTH
  * the largest per-channel absolute difference <= max_abs_diff (default 4)
  * the fraction of pixels with any differing channel <= max_diff_frac
    (default 0.01).

    python3 tests/golden/render_golden.py tests/golden/manifest.json             # check vs goldens
    python3 tests/golden/render_golden.py tests/golden/manifest.json --update    # (re)write goldens
    python3 tests/golden/render_golden.py tests/golden/manifest.json --selftest  # render twice, compare

Run from the repository root (commands and golden paths are relative to it).
Exit status 0 when every scene passes, 1 otherwise. A failing check writes the
full-size render next to the golden as <name>.actual.png.

Needs only numpy: PNG decode/encode (8-bit, non-interlaced) and the
resampling are implemented here, the resampling bit-compatible with Pillow's
Image.resize(..., Image.BILINEAR) on RGB images (separable triangle filter
whose support grows with the scale factor, 22-bit fixed-point coefficients,
horizontal pass first, 8-bit intermediate).
"""
from __future__ import annotations

import json
import os
import shlex
import shutil
import struct
import subprocess
import sys
import tempfile
import zlib

import numpy as np

DEFAULT_MAX_ABS_DIFF = 4
DEFAULT_MAX_DIFF_FRAC = 0.01
DOWNSCALE_WIDTH = 640
PRECISION_BITS = 32 - 8 - 2


# ----------------------------------------------------------------------- PNG
def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def _unfilter(raw: bytes, w: int, h: int, bpp: int) -> np.ndarray:
    stride = w * bpp
    out = np.zeros((h, stride), np.uint8)
    prev = np.zeros(stride, np.int32)
    pos = 0
    for y in range(h):
        ft = raw[pos]
        line = np.frombuffer(raw, np.uint8, stride, pos + 1).astype(np.int32)
        pos += stride + 1
        if ft == 0:
            cur = line
        elif ft == 1:   # Sub: running sum along each byte lane
            cur = (np.cumsum(line.reshape(w, bpp), axis=0) & 255).reshape(stride)
        elif ft == 2:   # Up
            cur = (line + prev) & 255
        elif ft == 3:   # Average
            cur = line.copy()
            for i in range(stride):
                left = cur[i - bpp] if i >= bpp else 0
                cur[i] = (cur[i] + ((left + prev[i]) >> 1)) & 255
        elif ft == 4:   # Paeth
            cur = line.copy()
            for i in range(stride):
                left = cur[i - bpp] if i >= bpp else 0
                upleft = prev[i - bpp] if i >= bpp else 0
                cur[i] = (cur[i] + _paeth(left, prev[i], upleft)) & 255
        else:
            raise ValueError(f"PNG: unknown filter type {ft} on row {y}")
        out[y] = cur
        prev = cur
    return out


def read_png(path) -> np.ndarray:
    """8-bit non-interlaced PNG -> (h, w, 3) uint8 RGB (alpha dropped, grey
    replicated, palette expanded: Image.convert('RGB'))."""
    with open(path, "rb") as f:
        data = f.read()
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError(f"{path}: not a PNG file")
    pos, idat, plte, hdr = 8, [], None, None
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        kind = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        pos += 12 + n
        if kind == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        elif kind == b"PLTE":
            plte = np.frombuffer(body, np.uint8).reshape(-1, 3)
        elif kind == b"IDAT":
            idat.append(body)
        elif kind == b"IEND":
            break
    if hdr is None:
        raise ValueError(f"{path}: no IHDR")
    w, h, depth, ctype, _, _, interlace = hdr
    if depth != 8 or interlace != 0:
        raise ValueError(f"{path}: only 8-bit non-interlaced PNGs are supported")
    bpp = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    px = _unfilter(zlib.decompress(b"".join(idat)), w, h, bpp).reshape(h, w, bpp)
    if ctype == 0 or ctype == 4:
        return np.repeat(px[..., :1], 3, axis=2)
    if ctype == 3:
        if plte is None:
            raise ValueError(f"{path}: palette image without PLTE")
        return plte[px[..., 0]]
    return np.ascontiguousarray(px[..., :3])


def write_png(path, rgb: np.ndarray):
    """(h, w, 3) uint8 -> 8-bit RGB PNG, filter type 0."""
    rgb = np.ascontiguousarray(rgb, np.uint8)
    h, w = rgb.shape[:2]
    raw = b"".join(b"\x00" + rgb[y].tobytes() for y in range(h))

    def chunk(kind, body):
        return struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body) & 0xFFFFFFFF)

    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0))
                + chunk(b"IDAT", zlib.compress(raw, 6)) + chunk(b"IEND", b""))


# ------------------------------------------------------------------ resample
def _bilinear_taps(in_size: int, out_size: int):
    """Per output index: first input index and the fixed-point weights of
    Pillow's bilinear (triangle) filter, support scaled by the reduction."""
    scale = in_size / out_size
    fscale = max(scale, 1.0)
    support = fscale
    ss = 1.0 / fscale
    taps = []
    for o in range(out_size):
        center = (o + 0.5) * scale
        lo = max(int(center - support + 0.5), 0)
        hi = min(int(center + support + 0.5), in_size)
        w = np.array([max(0.0, 1.0 - abs((x - center + 0.5) * ss)) for x in range(lo, hi)])
        s = w.sum()
        if s != 0.0:
            w = w / s
        k = np.where(w < 0, np.trunc(-0.5 + w * (1 << PRECISION_BITS)), np.trunc(0.5 + w * (1 << PRECISION_BITS)))
        taps.append((lo, k.astype(np.int64)))
    return taps


def _resample_axis1(img: np.ndarray, out_size: int) -> np.ndarray:
    """img (rows, in_size, ch) uint8 -> (rows, out_size, ch) uint8."""
    src = img.astype(np.int64)
    out = np.empty((img.shape[0], out_size, img.shape[2]), np.uint8)
    for o, (lo, k) in enumerate(_bilinear_taps(img.shape[1], out_size)):
        acc = (1 << (PRECISION_BITS - 1)) + np.einsum("rxc,x->rc", src[:, lo:lo + len(k)], k)
        out[:, o] = np.clip(acc >> PRECISION_BITS, 0, 255)
    return out


def resize_bilinear(rgb: np.ndarray, width: int, height: int) -> np.ndarray:
    out = rgb
    if width != rgb.shape[1]:
        out = _resample_axis1(out, width)
    if height != rgb.shape[0]:
        out = _resample_axis1(out.transpose(1, 0, 2), height).transpose(1, 0, 2)
    return np.ascontiguousarray(out)


def downscale(rgb: np.ndarray) -> np.ndarray:
    h, w = rgb.shape[:2]
    if w > DOWNSCALE_WIDTH:
        return resize_bilinear(rgb, DOWNSCALE_WIDTH, max(1, round(h * DOWNSCALE_WIDTH / w)))
    return rgb


# ------------------------------------------------------------------- compare
def load(path) -> np.ndarray:
    return downscale(read_png(path)).astype(np.int16)


def diff_stats(a, b):
    if a.shape != b.shape:
        return None
    d = np.abs(a.astype(np.int16) - b.astype(np.int16))
    px = np.any(d > 0, axis=2)
    return {"max_abs": int(d.max()), "diff_px": int(px.sum()), "diff_frac": float(px.mean()),
            "rmse": float(np.sqrt((d.astype(np.float64) ** 2).mean()))}


def verdict(s, max_abs_diff, max_diff_frac) -> bool:
    return s is not None and s["max_abs"] <= max_abs_diff and s["diff_frac"] <= max_diff_frac


def render(command: str, out_path: str):
    """Run a scene's command; None on success, else the reason it failed."""
    cmd = command.replace("{out}", out_path)
    r = subprocess.run(shlex.split(cmd), stdout=subprocess.DEVNULL, stderr=subprocess.PIPE)
    if r.returncode != 0:
        tail = r.stderr.decode(errors="replace").strip().splitlines()[-3:]
        return f"render command failed (exit {r.returncode}): {cmd}" + ("".join("\n        " + t for t in tail))
    if not os.path.exists(out_path):
        return f"render produced no output at {out_path}: {cmd}"
    return None


def run_manifest(manifest_path: str, mode: str = "check") -> bool:
    with open(manifest_path) as f:
        manifest = json.load(f)
    all_ok = True
    with tempfile.TemporaryDirectory() as tmp:
        for sc in manifest["scenes"]:
            name, golden = sc["name"], sc["golden"]
            mad = sc.get("max_abs_diff", DEFAULT_MAX_ABS_DIFF)
            mdf = sc.get("max_diff_frac", DEFAULT_MAX_DIFF_FRAC)
            actual = os.path.join(tmp, name + ".actual.png")
            err = render(sc["command"], actual)
            if mode == "--update":
                if err:
                    print(f"[update] {name}: FAILED — {err}")
                    all_ok = False
                    continue
                os.makedirs(os.path.dirname(golden) or ".", exist_ok=True)
                write_png(golden, downscale(read_png(actual)))
                print(f"[update] {name}: golden written -> {golden}")
                continue
            if err:
                print(f"[FAIL]  {name}: {err}")
                all_ok = False
                continue
            if mode == "--selftest":
                actual2 = os.path.join(tmp, name + ".actual2.png")
                err2 = render(sc["command"], actual2)
                if err2:
                    print(f"[FAIL]  {name}: {err2}")
                    all_ok = False
                    continue
                s = diff_stats(load(actual), load(actual2))
                ok = verdict(s, mad, mdf)
                print(f"[{'PASS' if ok else 'FAIL'}]  {name} (selftest): max_abs={s['max_abs']} "
                      f"diff_px={s['diff_px']} frac={s['diff_frac']:.6f} rmse={s['rmse']:.4f}")
                all_ok = all_ok and ok
                continue
            if not os.path.exists(golden):
                print(f"[FAIL]  {name}: golden missing ({golden}) — run with --update first")
                all_ok = False
                continue
            s = diff_stats(load(actual), load(golden))
            if s is None:
                print(f"[FAIL]  {name}: size mismatch vs golden")
                all_ok = False
                continue
            ok = verdict(s, mad, mdf)
            print(f"[{'PASS' if ok else 'FAIL'}]  {name}: max_abs={s['max_abs']}(<= {mad}) "
                  f"diff_frac={s['diff_frac']:.6f}(<= {mdf}) diff_px={s['diff_px']} rmse={s['rmse']:.4f}")
            if not ok:
                drift = golden[:-4] + ".actual.png" if golden.endswith(".png") else golden + ".actual.png"
                shutil.copyfile(actual, drift)
                print(f"        drift written -> {drift}")
            all_ok = all_ok and ok
    print("\nrender-golden:", "ALL PASS" if all_ok else "FAILURES — see above")
    return all_ok


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if not argv or argv[0].startswith("-") or (len(argv) > 1 and argv[1] not in ("--update", "--selftest")):
        print("usage: render_golden.py <manifest.json> [--update|--selftest]", file=sys.stderr)
        return 2
    return 0 if run_manifest(argv[0], argv[1] if len(argv) > 1 else "check") else 1


if __name__ == "__main__":
    sys.exit(main())
