"""Benchmark of the path-tracing core on one MI355X.

Workload: the procedural scene at scale — 64 displaced geodesic spheres
(32,000 triangles each) over a 1024x512-cell heightfield, 3,096,576 triangles,
GGX metallic-roughness materials, 2048x1024 sky + 0.5 degree sun environment,
1920x1080, 1 spp per step, max depth 4, NEE + MIS, PCG32 seed 7. One step =
one forward frame (generate -> extend/shade/connect x5 -> accumulate) with the
scene resident in HBM; the LBVH build is reported separately.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]

Prints one JSON line (rank 0) with the metric, a roofline entry for the
dominant kernel (closest-hit traversal, HBM/latency bound) and the CPU
baseline (the oracle's FP32 restatement on this host's cores, bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mpaths/s fwd & adjoint @1080p; GPU time to equal relMSE, denoised vs plain PT"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = ROOT / "profiles" / "extend_traffic_procedural_1080p.json"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def extend_bytes(rays, visits, tests):
    """Algorithmic bytes of the extend kernel: 32 B ray + 16 B hit per ray,
    64 B per node visit, 48 B per triangle test."""
    return 48.0 * rays + 64.0 * visits + 48.0 * tests


def cpu_baseline(sc, seconds, threads):
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib
    t0 = time.time()
    o = oracle_lib.OracleScene(sc)
    log(f"cpu baseline: oracle scene + LBVH in {time.time() - t0:.1f}s")
    # whole 1920x1080 frames at 1 spp (consecutive sample indices) until ~`seconds` of CPU work
    frames = 0
    t = time.time()
    while True:
        o.render(1, sc.settings.seed + frames, precision=32, threads=threads)
        frames += 1
        if time.time() - t >= seconds:
            break
    dt = time.time() - t
    paths = frames * sc.width * sc.height
    return {"value": paths / dt / 1e6, "unit": "Mpaths/s", "cores": threads, "kind": "port",
            "sample": f"{frames} full 1920x1080 frames at 1 spp, depth 4 ({paths} paths) of the bench workload, "
                      f"FP32 oracle restatement (same LBVH, traversal and integrator), {dt:.1f}s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from ohao_engine_amd import ohpt, scenes as S

    t0 = time.time()
    sc = S.procedural()
    log(f"scene: {sc.num_triangles} triangles generated in {time.time() - t0:.1f}s")
    ctx = ohpt.Context(sc.width, sc.height)
    t0 = time.time()
    ctx.load_scene(sc)
    build_ms = ctx.stats().ms_build
    log(f"upload + GPU LBVH: {time.time() - t0:.1f}s wall, LBVH build {build_ms:.2f} ms on GPU, "
        f"depth {ctx.tree_depth()}")

    spp = sc.settings.spp
    for i in range(a.warmup):
        ctx.render(spp, ohpt.RENDER_PROFILE)
    ctx.synchronize()
    log(f"warmup {a.warmup} steps done")

    acc = dict(ms_total=0.0, ms_extend=0.0, ms_shade=0.0, ms_connect=0.0, rays_extend=0, rays_shadow=0,
               node_visits=0, tri_tests=0, extend_launches=0)
    ctx.synchronize()
    t0 = time.perf_counter()
    for i in range(a.steps):
        ctx.render(spp, ohpt.RENDER_PROFILE)
        st = ctx.stats()
        for k in acc:
            acc[k] += getattr(st, k)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    log(f"timed {a.steps} steps: {wall * 1e3 / a.steps:.3f} ms/step wall, "
        f"{acc['ms_total'] / a.steps:.3f} ms/step GPU events")

    paths = sc.width * sc.height * spp * a.steps
    value = paths / wall / 1e6
    byts = extend_bytes(acc["rays_extend"], acc["node_visits"], acc["tri_tests"])
    achieved = byts / (acc["ms_extend"] * 1e-3) / 1e9 if acc["ms_extend"] > 0 else 0.0
    traffic, pmc = None, {}
    if TRAFFIC_FILE.exists():
        try:
            pmc = json.loads(TRAFFIC_FILE.read_text())
            traffic = pmc.get("bytes_per_launch")
        except Exception:
            traffic, pmc = None, {}
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "kernel": "k_extend (closest-hit LBVH traversal)",
                "launches": acc["extend_launches"],
                "avg_launch_ms": round(acc["ms_extend"] / max(acc["extend_launches"], 1), 5),
                "algorithmic_bytes_per_launch": round(byts / max(acc["extend_launches"], 1)),
                "node_visits_per_ray": round(acc["node_visits"] / max(acc["rays_extend"], 1), 3),
                "tri_tests_per_ray": round(acc["tri_tests"] / max(acc["rays_extend"], 1), 3),
                "l2_hit_pct": pmc.get("l2_hit_pct"), "occupancy_pct_first_bounce": pmc.get("occupancy_pct_first_bounce"),
                "traffic_source": "profiles/" + pmc["source"] if pmc.get("source") else None}
    breakdown = {k: round(acc[k] / a.steps, 4) for k in ("ms_total", "ms_extend", "ms_shade", "ms_connect")}
    breakdown["extend_rays_per_step"] = acc["rays_extend"] // a.steps
    breakdown["shadow_rays_per_step"] = acc["rays_shadow"] // a.steps

    # path-replay adjoint of the same frame: dL/dI = per-pixel weights, gradients
    # scattered into the 1024x512 ground albedo texture
    adj_steps = max(1, a.steps // 2)
    w = np.random.default_rng(0).uniform(0.5, 1.5, (sc.height, sc.width, 3)).astype(np.float32)
    ctx.zero_gradients()
    ctx.adjoint(w, spp)
    ctx.synchronize()
    t0 = time.perf_counter()
    adj_ms = 0.0
    for i in range(adj_steps):
        ctx.adjoint(w, spp)
        adj_ms += ctx.stats().ms_total
    adj_wall = time.perf_counter() - t0
    adjoint = {"value": round(sc.width * sc.height * spp * adj_steps / adj_wall / 1e6, 3), "unit": "Mpaths/s",
               "ms_per_step": round(adj_wall * 1e3 / adj_steps, 4),
               "gpu_ms_per_step": round(adj_ms / adj_steps, 4),
               "cost_vs_forward": round((adj_wall / adj_steps) / (wall / a.steps), 3),
               "parameters": "1024x512 RGB ground albedo texture", "mode": "fp32 atomics"}
    log(f"adjoint: {adjoint['ms_per_step']} ms/step, {adjoint['cost_vs_forward']}x forward")
    ctx.close()

    cpu = None
    if not a.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        try:
            cpu = cpu_baseline(sc, a.cpu_seconds, threads)
            log(f"cpu baseline: {cpu['value']:.3f} Mpaths/s on {threads} threads")
        except Exception as e:  # the baseline is reported, not required
            log(f"cpu baseline failed: {e}")

    out = {"metric": METRIC, "value": round(value, 3), "unit": "Mpaths/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": round(wall * 1e3 / a.steps, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "procedural: 64 displaced spheres x 32,000 + 1024x512 heightfield = "
                                  "3,096,576 triangles, 2048x1024 sky+sun env, 1920x1080, 1 spp, depth 4, "
                                  "NEE+MIS, PCG32 seed 7", "triangles": sc.num_triangles, "width": sc.width,
                      "height": sc.height, "spp_per_step": spp, "max_bounces": sc.settings.max_bounces,
                      "lbvh_build_ms": round(build_ms, 3)},
           "breakdown": breakdown, "adjoint": adjoint, "roofline": roofline, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
