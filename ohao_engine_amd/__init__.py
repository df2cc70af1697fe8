"""ohao_engine_amd — MI355X-native path-tracing core (libohpt.so, C-ABI in include/ohpt.h).

Modules:
  build     compile libohpt.so for gfx950 (hipcc)
  ohpt      ctypes binding + Context (mirror of the reference renderer calls)
  scenes    synthetic workloads (Cornell box, procedural 3.1M-triangle scene, lantern)
"""
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
__all__ = ["PACKAGE_DIR"]
