"""C++ façade and example drivers (ohao_engine_amd/bin/*), run as the
reference's drivers are run:
  * cornell_box through the façade renders the same image as the Python
    binding of the same scene, seed and settings (C1 at reduced size); its
    sample count follows the reference's render() x (spp + 3) convention;
  * turntable writes its frames; the learned denoise mode runs through the
    façade.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from ohao_engine_amd import ohpt, scenes as S

pytestmark = pytest.mark.gpu
BIN = Path(__file__).resolve().parents[1] / "ohao_engine_amd" / "bin"


def run(args, cwd, timeout=120):
    print("run:", " ".join(map(str, args)), flush=True)
    r = subprocess.run([str(a) for a in args], cwd=cwd, capture_output=True, text=True, timeout=timeout)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stderr
    return r.stdout


def read_pfm(path):
    with open(path, "rb") as f:
        assert f.readline().strip() == b"PF"
        w, h = map(int, f.readline().split())
        scale = float(f.readline())
        data = np.frombuffer(f.read(), "<f4" if scale < 0 else ">f4").reshape(h, w, 3)
    return data[::-1]


def test_cornell_box_driver_matches_binding(tmp_path):
    spp, size = 16, 64
    out = run([BIN / "cornell_box", tmp_path / "cb.png", spp, f"--size={size}x{size}", "--seed=1",
               f"--hdr={tmp_path / 'cb.pfm'}"], tmp_path)
    assert f"{spp} samples per pixel" in out
    png = (tmp_path / "cb.png").read_bytes()
    assert png[:8] == b"\x89PNG\r\n\x1a\n" and len(png) > size * size * 4
    img = read_pfm(tmp_path / "cb.pfm")
    sc = S.cornell_box(width=size, height=size, spp=spp, seed=1)
    ctx = ohpt.Context(size, size)
    ctx.load_scene(sc)
    ctx.render(spp)
    ref = ctx.read_accum()[..., :3]
    ctx.close()
    rel = float(np.abs(img - ref).sum() / np.abs(ref).sum())
    print(f"façade vs binding: mean abs rel diff {rel:.3e}, identical {np.array_equal(img, ref)}")
    assert rel < 1e-3


def test_turntable_driver_writes_frames(tmp_path):
    run([BIN / "turntable", "sphere", "cornell", 4, 2, "--size=96x64", f"--out={tmp_path / 'tt'}"], tmp_path)
    frames = sorted((tmp_path / "tt").glob("cornell_*.png"))
    assert len(frames) == 2 and all(f.stat().st_size > 96 * 64 * 4 for f in frames)


def test_learned_denoise_mode_through_facade(tmp_path):
    weights = Path(ohpt.__file__).resolve().parent / "weights" / "kpn_unet.bin"
    if not weights.exists():
        pytest.skip("no trained denoiser weights in the package")
    out = run([BIN / "cornell_box", tmp_path / "dn.png", 4, "--size=64x64", "--denoise=learned"], tmp_path)
    assert "Denoise mode: Learned" in out
    assert (tmp_path / "dn.png").stat().st_size > 64 * 64 * 4


def test_inverse_fit_driver_selftest(tmp_path):
    out = run([BIN / "inverse_fit", "--selftest"], tmp_path, timeout=300)
    assert "Done:" in out
