"""The C-ABI library loads and exports every entry point include/ohpt.h
declares; the Python binding covers them all; error paths return codes
without a GPU (no compute calls here)."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = []
    for h in sorted((ROOT / "include").glob("*.h")):
        txt = h.read_text()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(ohpt_[a-z0-9_]+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def test_header_declares_core_entry_points():
    names = declared_functions()
    for required in ["ohpt_create", "ohpt_destroy", "ohpt_scene_commit", "ohpt_render", "ohpt_read_accum",
                     "ohpt_trace_rays", "ohpt_last_error"]:
        assert required in names


def test_library_exports_every_declared_symbol(ohpt_lib):
    missing = [n for n in declared_functions() if not hasattr(ohpt_lib, n)]
    assert missing == []


def test_binding_covers_every_declared_symbol():
    from ohao_engine_amd import ohpt
    bound = {s[0] for s in ohpt.SIGNATURES}
    assert set(declared_functions()) <= bound


def test_error_paths_without_gpu(ohpt_lib):
    from ohao_engine_amd import ohpt
    L = ohpt_lib
    assert L.ohpt_abi_version() == 1
    # null context / arguments are rejected with a message, no device needed
    assert L.ohpt_render(None, 1, 0) == -1
    assert b"null" in L.ohpt_last_error()
    assert L.ohpt_destroy(None) == -1
    assert L.ohpt_set_camera(None, None) == -1
    assert L.ohpt_create(None, 0, 4, 4) == -1
    h = C.c_void_p()
    assert L.ohpt_create(C.byref(h), 0, 0, 4) == -1


def test_struct_layouts_match_header():
    from ohao_engine_amd import ohpt, scenes
    assert C.sizeof(ohpt.Camera) == 13 * 4
    assert C.sizeof(ohpt.Settings) == 32
    assert C.sizeof(ohpt.Stats) == 88
    assert scenes.MATERIAL_DTYPE.itemsize == 64


def test_denoiser_abi_without_gpu(ohpt_lib):
    from ohao_engine_amd import ohpt
    L = ohpt_lib
    assert C.sizeof(ohpt.DenoiserStats) == 32
    assert C.sizeof(ohpt.DenoiseInputs) == 4 * C.sizeof(C.c_void_p)
    assert L.ohpt_denoiser_create(None, 0, 0) == -1
    assert L.ohpt_denoiser_destroy(None) == 0
    assert L.ohpt_denoiser_infer(None, None, None, 1, 4, 4, 0) == -1
    assert L.ohpt_denoiser_adam_step(None, 1e-3) == -1
    assert L.ohpt_denoise(None, None, None, 0) == -1
    n = C.c_uint64()
    assert L.ohpt_denoiser_param_count(None, C.byref(n)) == -1


def test_facade_library_and_drivers_are_built():
    pkg = ROOT / "ohao_engine_amd"
    lib = C.CDLL(str(pkg / "libohao.so"))
    assert lib is not None
    for exe in ("cornell_box", "turntable", "inverse_fit"):
        assert (pkg / "bin" / exe).exists()
