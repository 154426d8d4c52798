"""Ray-cast microbenchmark: closest hit of 16M rays against a 1M-triangle soup.

Workload: 2^20 triangles, centroids uniform in the unit cube, vertex offsets
LogNormal(-5, 0.5) in uniform directions (seed-fixed); 16,777,216 rays =
8,388,608 coherent pinhole rays (4096 x 2048, eye (0.5, 0.5, -1.5), FOV 40)
+ 8,388,608 incoherent rays (origins uniform in the cube, directions uniform
on the sphere).  A step is one closest-hit pass over all 16M rays with the
rays already resident in HBM; the BVH is built once before timing.

Prints one JSON line (metric Mrays/s, higher is better) with the dominant
kernel's roofline (HBM-bound: algorithmic bytes = 32 B ray + 16 B hit + 128 B
per BVH4 node fetched + 40 B per triangle tested, node/triangle counts from
the kernel's own counters) and the CPU oracle's BVH traversal timed on the
host cores on a bounded sample of the same rays.
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

HBM_PEAK_GBS = 8000.0
PMC_SUMMARY = ROOT / "profiles" / "trace_raycast_1M_16M_pmc.json"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--triangles", type=int, default=1 << 20)
    ap.add_argument("--width", type=int, default=4096)
    ap.add_argument("--height", type=int, default=2048)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-halves", action="store_true", help="skip the per-half (coherent / incoherent) timing")
    ap.add_argument("--leaf", type=int, default=4, help="BVH leaf size (1..16)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 512 MiB device copy (reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (rays and pixels would be sharded "
              "by range with the BVH replicated per GPU; see DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        return 2

    import torch

    from master_thesis_amd import scenes
    from master_thesis_amd._lib import RAY_DTYPE
    from master_thesis_amd.tracer import Tracer

    t0 = time.time()
    sc = scenes.triangle_soup(args.triangles, seed=1234)
    coh, inc = scenes.microbench_rays(args.width, args.height, seed=4321)
    rays = np.concatenate([coh, inc])
    n = rays.shape[0]
    n_coh = coh.shape[0]
    log(f"[bench] scene {sc.num_triangles} tris, {n} rays generated in {time.time() - t0:.1f}s")

    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(dev)  # one explicit stream for the library's launches and the timing events
    torch.cuda.set_stream(stream)
    d_rays = torch.from_numpy(rays.view(np.float32).reshape(-1)).to(dev)
    d_hits = torch.empty(n * 4, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    if args.calibrate_copy:
        src = torch.empty(1 << 27, dtype=torch.float32, device=dev)  # 512 MiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("[bench] calibration: 3 x 512 MiB device copy done")
    with Tracer(0) as t:
        t.set_stream(stream.cuda_stream)
        t.upload(sc)
        info = t.build(args.leaf)
        info = t.build(args.leaf)  # second build: steady-state build time
        log(f"[bench] BVH4: {info['num_nodes']} nodes, depth {info['depth']}, build {info['build_ms']:.2f} ms")

        # counters pass (same kernel with per-wave counters): node / triangle fetches for the roofline
        _, cnt = t.trace(d_rays, d_hits, counters=True)
        torch.cuda.synchronize()
        bytes_alg = 32 * cnt["rays"] + 16 * cnt["rays"] + 128 * cnt["node_visits"] + 40 * cnt["tri_tests"]

        # per-half rates (informational)
        halves = {"coherent": None, "incoherent": None}
        for name, lo, hi in (("coherent", 0, n_coh), ("incoherent", n_coh, n)) if not args.no_halves else ():
            sub = d_rays[lo * 8:hi * 8]
            sub_h = d_hits[lo * 4:hi * 4]
            for _ in range(2):
                t.trace(sub, sub_h)
            ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            ev0.record(stream)
            for _ in range(5):
                t.trace(sub, sub_h)
            ev1.record(stream)
            torch.cuda.synchronize()
            halves[name] = (hi - lo) * 5 / (ev0.elapsed_time(ev1) / 1e3) / 1e6
        log(f"[bench] per half Mrays/s: {halves}")

        for _ in range(args.warmup):
            t.trace(d_rays, d_hits)
        torch.cuda.synchronize()
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        kernel_ms = []
        torch.cuda.synchronize()
        start.record(stream)
        for _ in range(args.steps):
            t.trace(d_rays, d_hits)
        end.record(stream)
        torch.cuda.synchronize()
        total_ms = start.elapsed_time(end)
        # kernel-only duration of the traversal launch (HIP events around the kernel itself)
        for _ in range(3):
            t.trace(d_rays, d_hits)
            kernel_ms.append(t.timing()["last_kernel_ms"])
        hits = d_hits.cpu().numpy().view(np.uint32).reshape(-1, 4)

    ms_per_step = total_ms / args.steps
    value = n / (ms_per_step / 1e3) / 1e6
    k_ms = float(np.mean(kernel_ms))
    achieved = bytes_alg / (k_ms / 1e3) / 1e9
    traffic = None
    if PMC_SUMMARY.exists():
        try:
            traffic = json.loads(PMC_SUMMARY.read_text()).get("hbm_bytes_per_launch")
        except (ValueError, OSError):
            traffic = None
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "k_trace", "kernel_ms": round(k_ms, 4),
                "bytes_alg_per_launch": int(bytes_alg),
                "node_visits_per_ray": round(cnt["node_visits"] / n, 3),
                "tri_tests_per_ray": round(cnt["tri_tests"] / n, 3)}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(sc, rays, hits, args.cpu_seconds)

    out = {
        "metric": "Mrays/s closest-hit",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"ray-cast microbench: {sc.num_triangles}-triangle lognormal soup, "
                               f"{n} rays ({n_coh} coherent pinhole + {n - n_coh} incoherent), closest hit",
                   "triangles": sc.num_triangles, "rays": n, "bvh": f"BVH4 LBVH leaf<={args.leaf}",
                   "bvh_build_ms": round(info["build_ms"], 3),
                   "mrays_per_s_coherent": halves["coherent"] and round(halves["coherent"], 1),
                   "mrays_per_s_incoherent": halves["incoherent"] and round(halves["incoherent"], 1),
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def cpu_baseline(sc, rays, gpu_hits, target_s: float):
    """The oracle's CPU BVH (binned SAH) on a bounded, evenly strided sample
    of the same rays, all host threads up to 16; also checks the GPU hits on
    that sample."""
    import oracle as O
    threads = min(16, os.cpu_count() or 1)
    t0 = time.time()
    bvh = O.CpuBVH(sc)
    build_s = time.time() - t0
    probe = rays[:: max(1, rays.shape[0] // 65536)].copy()
    t0 = time.time()
    bvh.trace(probe, threads)
    per_ray = (time.time() - t0) / probe.shape[0]
    m = int(min(rays.shape[0], max(65536, target_s / max(per_ray, 1e-12))))
    stride = max(1, rays.shape[0] // m)
    idx = np.arange(0, rays.shape[0], stride)[:m]
    sample = rays[idx].copy()
    t0 = time.time()
    ch, _ = bvh.trace(sample, threads)
    dt = time.time() - t0
    same = bool(np.array_equal(ch.view(np.uint32).reshape(-1, 4), gpu_hits[idx]))
    log(f"[bench] CPU oracle: {sample.shape[0]} rays in {dt:.2f}s on {threads} threads (BVH build {build_s:.1f}s); "
        f"hits identical to GPU: {same}")
    return {"value": round(sample.shape[0] / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"every {stride}th ray of the same 16M ({sample.shape[0]} rays, both halves), oracle binned-SAH "
                      f"BVH built in {build_s:.1f}s; hits identical to GPU: {same}"}


if __name__ == "__main__":
    sys.exit(main())
