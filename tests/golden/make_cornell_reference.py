"""Writes tests/golden/cornell_fp64_4096spp.npz: the FP64-accumulated oracle
render of the 256x256 Cornell box (max depth 8, RR from bounce 3, PCG32) at
4096 spp with sample indices starting at 1,000,000 (independent of the GPU
runs, which start at seed 1). Stores the per-pixel mean (float32) and the
per-pixel sample variance (float32).

    python tests/golden/make_cornell_reference.py [--spp 4096] [--threads 8]
"""
import argparse
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import oracle_lib  # noqa: E402
from ohao_engine_amd import scenes as S  # noqa: E402

FIRST_SAMPLE = 1_000_000


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--spp", type=int, default=4096)
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--out", default=str(Path(__file__).with_name("cornell_fp64_4096spp.npz")))
    a = ap.parse_args()
    sc = S.cornell_box()
    o = oracle_lib.OracleScene(sc)
    t0 = time.time()
    s = np.zeros((sc.height, sc.width, 3))
    q = np.zeros_like(s)
    chunk = 256
    done = 0
    while done < a.spp:
        k = min(chunk, a.spp - done)
        ss, qq = o.render(k, FIRST_SAMPLE + done, precision=64, threads=a.threads)
        s += ss
        q += qq
        done += k
        print(f"{done}/{a.spp} spp  {time.time() - t0:.0f}s", flush=True)
    mean = s / a.spp
    var = np.maximum(q / a.spp - mean ** 2, 0.0) * a.spp / (a.spp - 1)
    np.savez_compressed(a.out, mean=mean.astype(np.float32), var=var.astype(np.float32),
                        spp=np.int64(a.spp), first_sample=np.int64(FIRST_SAMPLE))
    print(f"wrote {a.out} in {time.time() - t0:.0f}s")


if __name__ == "__main__":
    main()
