"""Ray-cast microbenchmark: closest hit of 16M rays against a 1M-triangle soup.

Workload: 2^20 triangles, centroids uniform in the unit cube, vertex offsets
LogNormal(-5, 0.5) in uniform directions (seed-fixed); 16,777,216 rays =
8,388,608 coherent pinhole rays (4096 x 2048, eye (0.5, 0.5, -1.5), FOV 40)
+ 8,388,608 incoherent rays (origins uniform in the cube, directions uniform
on the sphere).  A step is one closest-hit pass over all 16M rays with the
rays already resident in HBM; the BVH is built once before timing.

Prints one JSON line (metric Mrays/s, higher is better) with the dominant
kernel's roofline (HBM: algorithmic bytes = 32 B ray + 16 B hit + 64 B per
BVH4 traversal node fetched + 40 B per triangle tested, node/triangle counts
from the kernel's own counters; beside them the memory-side traffic, L2 hit
rate and VALU issue rate that rocprofv3 counters measured on the same
kernel sources) and the CPU oracle's BVH traversal timed on the host cores on
a bounded sample of the same rays.
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
NODE_BYTES = 64  # rt_bvh4q_node, the traversal's node
PMC_SUMMARY = ROOT / "profiles" / "trace_raycast_1M_16M_pmc.json"
# the sources that define the traversal kernel and the tree it walks: counter
# summaries record their hash, and figures from other sources are not used
KERNEL_SOURCES = ["include/rt.h", "master_thesis_amd/csrc/rt_core.hpp", "master_thesis_amd/csrc/rt_internal.hpp",
                  "master_thesis_amd/csrc/trace.hip", "master_thesis_amd/csrc/lbvh.hip"]


def kernel_source_hash() -> str:
    import hashlib
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        h.update((ROOT / f).read_bytes())
    return h.hexdigest()[:16]


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def host_threads() -> tuple[int, dict]:
    """Worker threads for a CPU baseline: every CPU this process may run on
    (its affinity mask and cgroup CPU quota), within OMP_NUM_THREADS when the
    environment sets it (the GPU box grants one GPU's share of its host cores
    that way).  Returns (threads, a description of the host)."""
    logical = os.cpu_count() or 1
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else logical
    limits = []
    try:
        quota, period = Path("/sys/fs/cgroup/cpu.max").read_text().split()[:2]
        if quota != "max":
            q = max(1, int(-(-int(quota) // int(period))))
            if q < n:
                n, limits = q, limits + [f"cgroup quota {q}"]
    except (OSError, ValueError):
        pass
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and 0 < int(omp) < n:
        n, limits = int(omp), limits + [f"OMP_NUM_THREADS={omp}"]
    model, tpc = "", None
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name") and not model:
                model = line.split(":", 1)[1].strip()
        sib = Path("/sys/devices/system/cpu/cpu0/topology/thread_siblings_list").read_text().strip()
        tpc = len([x for part in sib.split(",") for x in (range(int(part.split("-")[0]), int(part.split("-")[-1]) + 1))])
    except (OSError, ValueError):
        pass
    return n, {"cpu_model": model, "logical_cpus": logical, "threads_per_core": tpc,
               "threads_used": n, "limited_by": ", ".join(limits) or "none"}


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
    ap.add_argument("--config", default="raycast",
                    choices=["raycast", "terrain_raycast", "cornell", "terrain", "diff", "nrc", "bdpt", "denoise",
                             "mlp"],
                    help="raycast: the headline microbench (default); cornell / terrain / diff / nrc: the other "
                         "workloads")
    ap.add_argument("--nrc-scene", default="terrain",
                    choices=["terrain", "cornell_closed", "cornell_indirect", "cornell_box"],
                    help="nrc: the scene")
    ap.add_argument("--frames", type=int, default=300, help="nrc: online training frames")
    ap.add_argument("--ref-spp", type=int, default=4096, help="nrc: samples of the light-sampled reference")
    ap.add_argument("--res", default="1920x1080", help="nrc: image size")
    ap.add_argument("--views", type=int, default=16, help="terrain: camera views batched in one launch")
    ap.add_argument("--spp", type=int, default=0, help="cornell / terrain / diff: override samples per pixel")
    ap.add_argument("--summarize-pmc", metavar="DIR", default="",
                    help="raycast / terrain_raycast: write the counter summaries under profiles/ from the "
                         "rocprofv3 CSV files of the passes in PMC_PASSES under DIR (no GPU work)")
    ap.add_argument("--summarize-stats", nargs=3, metavar=("DIR", "NAME", "CMD"), default=None,
                    help="write profiles/NAME_kernel_stats.md from the rocprofv3 --kernel-trace --stats CSV files "
                         "under DIR that the command CMD wrote (no GPU work)")
    args = ap.parse_args()
    if args.summarize_stats:
        summarize_stats(Path(args.summarize_stats[0]), *args.summarize_stats[1:])
        return 0
    if args.summarize_pmc:
        summarize_pmc(args.config, Path(args.summarize_pmc))
        return 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (rays and pixels would be sharded "
              "by range with the BVH replicated on every GPU). Run with --gpus 1.", file=sys.stderr)
        return 2
    if args.config not in ("raycast", "terrain_raycast"):
        out = {"cornell": bench_cornell, "terrain": bench_terrain, "diff": bench_diff,
               "nrc": bench_nrc, "bdpt": bench_bdpt, "denoise": bench_denoise, "mlp": bench_mlp}[args.config](args)
        print(json.dumps(out), flush=True)
        return 0
    if args.config == "terrain_raycast":
        sc, rays, n_coh = terrain_rays()
        workload = (f"ray-cast on the terrain: {sc.num_triangles} triangles (fBm heightfield 2048x2048 + 64 "
                    f"icospheres), {rays.shape[0]} rays ({n_coh} primary rays of 4 views at 1920x1080 + "
                    f"{rays.shape[0] - n_coh} first diffuse bounces from their hits), closest hit")
        out = raycast(args, sc, rays, n_coh, workload, ROOT / "profiles" / "trace_terrain_16M_pmc.json", cluster=True)
        out["metric"] = "Grays/s closest-hit (9.69M-triangle terrain)"
        out["value"] = round(out["value"] / 1e3, 4)
        out["unit"] = "Grays/s"
        print(json.dumps(out), flush=True)
        return 0

    from master_thesis_amd import scenes

    t0 = time.time()
    sc = scenes.triangle_soup(args.triangles, seed=1234)
    coh, inc = scenes.microbench_rays(args.width, args.height, seed=4321)
    rays = np.concatenate([coh, inc])
    n_coh = coh.shape[0]
    log(f"[bench] scene {sc.num_triangles} tris, {rays.shape[0]} rays generated in {time.time() - t0:.1f}s")
    workload = (f"ray-cast microbench: {sc.num_triangles}-triangle lognormal soup, {rays.shape[0]} rays "
                f"({n_coh} coherent pinhole + {rays.shape[0] - n_coh} incoherent), closest hit")
    print(json.dumps(raycast(args, sc, rays, n_coh, workload, PMC_SUMMARY)), flush=True)
    return 0


def terrain_rays(views: int = 4, width: int = 1920, height: int = 1080, seed: int = 77):
    """Rays of the terrain workload: the primary rays of `views` cameras on
    the terrain circle, then as many first diffuse bounces (uniform on the
    hemisphere of the shading normal, origin offset 0.01) from their hits,
    the hits drawn with replacement."""
    from master_thesis_amd import scenes
    from master_thesis_amd._lib import RAY_DTYPE
    from master_thesis_amd.tracer import Tracer
    t0 = time.time()
    sc = scenes.terrain_scene()
    prim = []
    for v in range(views):
        a = 2 * np.pi * v / views
        eye = (42.0 * np.cos(a), 22.0, 42.0 * np.sin(a))
        prim.append(scenes.pinhole_rays(eye, (0.0, 0.0, 0.0), 45.0, width, height))
    prim = np.concatenate(prim)
    with Tracer(0) as t:
        t.upload(sc)
        t.build(2)
        h = t.trace(prim)
    ok = h["tri"] != 0xFFFFFFFF
    rng = np.random.default_rng(seed)
    pick = rng.choice(np.nonzero(ok)[0], prim.shape[0])
    o = prim["o"][pick].astype(np.float64)
    d = prim["d"][pick].astype(np.float64)
    p = o + h["t"][pick][:, None] * d
    tri = h["tri"][pick]
    u, v = h["u"][pick][:, None], h["v"][pick][:, None]
    nv = sc.nrm[sc.idx[tri]].astype(np.float64)  # (n, 3 vertices, 3)
    n = (1 - u - v) * nv[:, 0] + u * nv[:, 1] + v * nv[:, 2]
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    n *= np.where(np.sum(n * d, axis=1) > 0, -1.0, 1.0)[:, None]  # face the incoming ray
    w = rng.normal(size=n.shape)
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    w *= np.where(np.sum(w * n, axis=1) < 0, -1.0, 1.0)[:, None]
    sec = np.zeros(prim.shape[0], RAY_DTYPE)
    sec["o"] = (p + 0.01 * w).astype(np.float32)
    sec["d"] = w.astype(np.float32)
    sec["tmax"] = np.inf
    log(f"[bench] terrain {sc.num_triangles} tris, {2 * prim.shape[0]} rays ({ok.mean():.3f} of the primary rays "
        f"hit) generated in {time.time() - t0:.1f}s")
    return sc, np.concatenate([prim, sec]), prim.shape[0]


def raycast(args, sc, rays, n_coh, workload, pmc_path, cluster=False):
    """Closest hit of all rays, resident in HBM, one step = one pass; the
    JSON record of the measurement.  cluster: the clustered hierarchy
    (RT_BUILD_CLUSTER) in place of the Morton one."""
    import torch

    from master_thesis_amd.tracer import Tracer
    n = rays.shape[0]
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
        info = t.build(args.leaf, cluster=cluster)
        info = t.build(args.leaf, cluster=cluster)  # second build: steady-state build time
        log(f"[bench] BVH4: {info['num_nodes']} nodes, depth {info['depth']}, build {info['build_ms']:.2f} ms")

        # counters pass (same kernel with per-wave counters): node / triangle fetches for the roofline
        _, cnt = t.trace(d_rays, d_hits, counters=True)
        torch.cuda.synchronize()
        bytes_alg = 32 * cnt["rays"] + 16 * cnt["rays"] + NODE_BYTES * cnt["node_visits"] + 40 * cnt["tri_tests"]

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
    # counter summary (profiles/): used only when it was measured on these kernel sources
    pmc, src = None, kernel_source_hash()
    if pmc_path.exists():
        try:
            pmc = json.loads(pmc_path.read_text())
        except (ValueError, OSError):
            pmc = None
    same = pmc is not None and pmc.get("source_hash") == src
    traffic = pmc.get("hbm_bytes_per_launch") if same else None
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                # memory-side bytes (L2 misses to the fabric, MALL and HBM; FETCH_SIZE + WRITE_SIZE,
                # calibrated) over this run's kernel time, against HBM's peak
                "traffic_gbps": traffic and round(traffic / (k_ms / 1e3) / 1e9, 1),
                "traffic_frac": traffic and round(traffic / (k_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                "l2_hit_rate": pmc.get("l2_hit_rate") if same else None,
                # wave64 VALU instructions per CU per cycle: a wave64 op that occupies its SIMD for 4
                # cycles allows 1.0; measured values at or above 1 mean the VALU pipes are saturated
                "valu_insts_per_cu_cycle": pmc.get("valu_insts_per_cu_cycle") if same else None,
                # rocprof's VALUBusy for gfx950: SQ_ACTIVE_INST_VALU / CUs / GRBM_GUI_ACTIVE cycles
                "valu_busy": pmc.get("valu_busy") if same else None,
                "counters": (pmc_path.relative_to(ROOT).as_posix() if same else
                             f"none for these kernel sources ({src})"),
                "kernel": "k_trace", "kernel_ms": round(k_ms, 4),
                "bytes_alg_per_launch": int(bytes_alg), "node_bytes": NODE_BYTES,
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
        "config": {"workload": workload,
                   "triangles": sc.num_triangles, "rays": n,
                   "bvh": f"BVH4 from the {'clustered (PLOC)' if cluster else 'Morton-order (LBVH)'} binary tree, "
                          f"SAH leaves of <= {args.leaf}",
                   "bvh_build_ms": round(info["build_ms"], 3),
                   # the build's algorithmic bytes per triangle: its 40-byte record read and an 8-byte
                   # key / id written, 4 radix passes reading and writing key and value (16 B each),
                   # ~64 B of hierarchy, boxes and collapse
                   "bvh_build_gbps_alg": round(sc.num_triangles * (48 + 4 * 16 + 64) / (info["build_ms"] * 1e-3) / 1e9, 1),
                   "mrays_per_s_coherent": halves["coherent"] and round(halves["coherent"], 1),
                   "mrays_per_s_incoherent": halves["incoherent"] and round(halves["incoherent"], 1),
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    return out


def cpu_baseline(sc, rays, gpu_hits, target_s: float):
    """The oracle's CPU BVH (binned SAH) on a bounded, evenly strided sample
    of the same rays, all host threads up to 16; also checks the GPU hits on
    that sample."""
    import oracle as O
    threads, host = host_threads()
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
    return {"value": round(sample.shape[0] / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "host": host, "kind": "port",
            "sample": (f"all {rays.shape[0]} rays" if stride == 1 else
                       f"every {stride}th ray of the same {rays.shape[0]} ({sample.shape[0]} rays, both halves)") +
                      f", oracle binned-SAH "
                      f"BVH built in {build_s:.1f}s (a fast CPU port: the reference's own agglomerative build is "
                      f"O(n^3) and cannot build 1M triangles); hits identical to GPU: {same}"}


def _render_timed(t, cams, s, reps: int):
    """Median over reps of the library's own device time for one render call."""
    times, rays = [], 0
    for _ in range(reps):
        rgb, _ = t.render_views(cams, s)
        tm = t.timing()
        times.append(tm["last_call_ms"])
        rays = tm["rays_traced"]
    return rgb, float(np.median(times)), rays


def bench_cornell(args):
    """Cornell-style box, 256 x 256, 16 spp, depth 5 (the reference estimator):
    spp/s and Mrays/s on the GPU, relative RMSE against a 4096-spp render
    (the deterministic GPU mode, which equals the oracle bit for bit), and the
    oracle on the host cores over the same frame."""
    from master_thesis_amd import scenes, host_lib as H
    from master_thesis_amd._lib import RT_RENDER_DETERMINISTIC
    from master_thesis_amd.tracer import Tracer, make_settings
    spp = args.spp or 16
    sc = scenes.cornell_box()
    cam = H.cornell_camera(256, 256)
    s = make_settings(spp=spp, max_depth=5, seed=3)
    with Tracer(0) as t:
        t.upload(sc)
        t.build(args.leaf)
        _render_timed(t, [cam], s, 2)
        rgb, ms, rays = _render_timed(t, [cam], s, 5)
        ref, _ = t.render(cam, make_settings(spp=4096, max_depth=5, seed=99, flags=RT_RENDER_DETERMINISTIC))
    rel_rmse = float(np.sqrt(np.mean((rgb[0] - ref) ** 2)) / np.sqrt(np.mean(ref ** 2)))
    # the same on the saved frames: the reference's clamp and 8-bit truncation
    from master_thesis_amd.images import to_rgb8
    a8, r8 = to_rgb8(rgb[0]).astype(np.float64), to_rgb8(ref).astype(np.float64)
    rel_rmse8 = float(np.sqrt(np.mean((a8 - r8) ** 2)) / np.sqrt(np.mean(r8 ** 2)))
    P = 256 * 256
    cpu = None
    if not args.no_cpu:
        import oracle as O
        threads, host = host_threads()
        # the reference renderer in faithful mode: its agglomerative two-level
        # BVHs, enumerate-all queries, one mutex-guarded RNG, one pixel per
        # atomic fetch_add; a bounded run of pixels when the frame would take
        # longer than --cpu-seconds
        _, _, probe_s = O.render_faithful(sc, cam, spp, 5, threads=threads, seed=3, pixel_limit=4096)
        lim = int(min(P, max(4096, 4096 * args.cpu_seconds / max(probe_s, 1e-9))))
        _, _, dt = O.render_faithful(sc, cam, spp, 5, threads=threads, seed=3, pixel_limit=lim)
        t0 = time.time()
        O.render(sc, cam, s, threads)
        dt_port = time.time() - t0
        cpu = {"value": round(lim * spp / dt, 1), "unit": "spp/s", "cores": threads, "host": host, "kind": "port",
               "sample": (f"the reference renderer restated in faithful mode (oracle/faithful.cpp: agglomerative "
                          f"two-level BVHs, enumerate-all queries, one mutex-guarded mt19937, one pixel per atomic "
                          f"fetch_add) over the first {lim} pixels of the frame at {spp} spp, depth 5; "
                          f"its time on this frame is 0.90x (1 thread) to 1.04x (8 threads) the compiled "
                          f"reference's, measured beside it in the build container"),
               "port_binned_sah": {"value": round(P * spp / dt_port, 1), "unit": "spp/s",
                                   "sample": "the whole frame with the oracle's binned-SAH BVH and Philox "
                                             "(a fast CPU port, not the reference's cost structure)"}}
    return {"metric": "spp/s (Cornell-style box, 256x256, 16 spp, depth 5)", "value": round(P * spp / (ms / 1e3), 1),
            "unit": "spp/s", "ms_per_frame": round(ms, 3), "mrays_per_s": round(rays / (ms / 1e3) / 1e6, 1),
            "rays_per_frame": int(rays), "rel_rmse_vs_4096spp": round(rel_rmse, 4),
            "rel_rmse_vs_4096spp_8bit": round(rel_rmse8, 4), "higher_is_better": True,
            "dtype": "f32", "data": "synthetic", "n_gpus": 1,
            "config": {"workload": "Cornell-style box: 5 walls, 2 boxes, 1 light (36 triangles), FOV 40, 256x256, "
                                   f"{spp} spp, depth 5", "triangles": int(sc.num_triangles)},
            "cpu_baseline": cpu}


def bench_bdpt(args):
    """The bidirectional tracer on the Cornell-style box, 256 x 256, with the
    reference's SoftwareRendererBidirectionalPathTracer settings (vision 5,
    light 5, maxDepth 5; 16 spp here): spp/s on the GPU, the oracle's
    restatement on the host cores over a 64 x 64 crop of the same frame."""
    from master_thesis_amd import scenes, host_lib as H
    from master_thesis_amd.tracer import Tracer, make_settings
    spp = args.spp or 16
    sc = scenes.cornell_box()
    cam = H.cornell_camera(256, 256)
    with Tracer(0) as t:
        t.upload(sc)
        t.build(args.leaf)
        t.render_bdpt(cam, spp, 5, 5, 5, seed=3)
        ms = []
        for r in range(5):
            t.render_bdpt(cam, spp, 5, 5, 5, seed=3, sample_offset=spp * r)
            ms.append(t.timing()["last_call_ms"])
        ms = float(np.median(ms))
    P = 256 * 256
    cpu = None
    if not args.no_cpu:
        import oracle as O
        threads, host = host_threads()
        crop = H.cornell_camera(64, 64)
        t0 = time.time()
        O.render_bdpt(sc, crop, make_settings(spp=spp, max_depth=5, seed=3), 5, 5, 5, threads)
        dt = time.time() - t0
        cpu = {"value": round(64 * 64 * spp / dt, 1), "unit": "spp/s", "cores": threads, "host": host, "kind": "port",
               "sample": f"a 64x64 view of the same box at {spp} spp with the oracle's CPU BVH"}
    return {"metric": f"spp/s (bidirectional tracer, Cornell-style box, 256x256, {spp} spp, 5/5/5)",
            "value": round(P * spp / (ms / 1e3), 1), "unit": "spp/s", "ms_per_frame": round(ms, 3),
            "higher_is_better": True, "dtype": "f32", "data": "synthetic", "n_gpus": 1,
            "config": {"workload": f"Cornell-style box (36 triangles), FOV 40, 256x256, {spp} spp, visionJumpCount 5, "
                                   "lightJumpCount 5, maxDepth 5", "triangles": int(sc.num_triangles)},
            "cpu_baseline": cpu}


def bench_terrain(args):
    """fBm terrain (2048^2, 8.38M triangles) + 64 icospheres (1.31M), sun and
    sky, 1920x1080, 64 spp, depth 8, 16 views in one launch sequence."""
    from master_thesis_amd import scenes
    from master_thesis_amd._lib import RT_RENDER_SKY, RT_RENDER_SUN
    from master_thesis_amd.tracer import Tracer, make_settings
    spp = args.spp or 64
    t0 = time.time()
    sc = scenes.terrain_scene()
    cams = scenes.terrain_cameras(args.views, 1920, 1080)
    log(f"[bench] terrain {sc.num_triangles} tris generated in {time.time() - t0:.1f}s")
    s = make_settings(spp=spp, max_depth=8, seed=5, flags=RT_RENDER_SKY | RT_RENDER_SUN, sky=(0.45, 0.55, 0.75),
                      sun_dir=(0.4, 0.8, 0.3), sun_radius_deg=5.0, sun_radiance=(40.0, 36.0, 30.0))
    with Tracer(0) as t:
        t.upload(sc)
        info = t.build(args.leaf, cluster=True)
        info = t.build(args.leaf, cluster=True)
        log(f"[bench] BVH4 {info['num_nodes']} nodes depth {info['depth']} build {info['build_ms']:.1f} ms")
        rgb, ms, rays = _render_timed(t, cams, s, 1)
    P = 1920 * 1080 * len(cams)
    cpu = None
    if not args.no_cpu:
        import oracle as O
        from master_thesis_amd import host_lib as H  # noqa: F401
        threads, host = host_threads()
        one = scenes.terrain_cameras(1, 1920, 1080)
        s1 = make_settings(spp=4, max_depth=8, seed=5, flags=s.flags, sky=(0.45, 0.55, 0.75),
                           sun_dir=(0.4, 0.8, 0.3), sun_radius_deg=5.0, sun_radiance=(40.0, 36.0, 30.0))
        t0 = time.time()
        O.render(sc, one[0], s1, threads)
        dt = O.last_render_seconds()
        log(f"[bench] terrain oracle: {time.time() - t0:.1f}s with its BVH build, {dt:.1f}s path tracing")
        cpu = {"value": round(1920 * 1080 * 4 / dt, 1), "unit": "spp/s", "cores": threads, "host": host, "kind": "port",
               "sample": "view 0 at 1920x1080, 4 spp, depth 8 (the oracle traces every sample from the camera); "
                         "path-tracing time only, its CPU BVH build over 9.69M triangles excluded; a fast CPU port "
                         "(binned-SAH BVH, Philox): the reference's agglomerative O(n^3) build cannot build this "
                         "scene, so its faithful mode cannot run here"}
    return {"metric": f"spp/s path tracing (terrain, {len(cams)} views at 1920x1080, {spp} spp, depth 8)",
            "value": round(P * spp / (ms / 1e3), 1), "unit": "spp/s",
            "grays_per_s": round(rays / (ms / 1e3) / 1e9, 3), "rays": int(rays),
            "rays_per_sample": round(rays / (P * spp), 3),
            "ms_per_batch": round(ms, 1), "views": len(cams), "higher_is_better": True,
            "dtype": "f32", "data": "synthetic", "n_gpus": 1, "bvh_build_ms": round(info["build_ms"], 2),
            "config": {"workload": "fBm heightfield 2048x2048 (6 octaves) + 64 level-5 icospheres, sun + sky, "
                                   f"1920x1080, {spp} spp, depth 8, {len(cams)} views in one launch sequence; "
                                   "each pixel's fixed primary ray traced once per work item",
                       "triangles": int(sc.num_triangles)},
            "image_mean": [round(float(x), 4) for x in rgb.reshape(-1, 3).mean(axis=0)],
            "cpu_baseline": cpu}


def bench_diff(args):
    """Differentiable Cornell-style box, 64 albedo groups, 256x256, 32 spp,
    depth 4: time of forward render + path-replay adjoint, agreement with
    central differences (h = 1e-3, common random numbers), oracle adjoint on
    the host cores."""
    from master_thesis_amd import scenes, host_lib as H
    from master_thesis_amd.tracer import Tracer, make_settings
    spp = args.spp or 32
    sc = scenes.cornell_groups(seed=0)
    cam = H.cornell_camera(256, 256)
    s = make_settings(spp=spp, max_depth=4, seed=7)
    target_sc = scenes.cornell_groups(seed=1)
    w = None
    with Tracer(0) as t:
        t.upload(target_sc)
        t.build(args.leaf)
        target, _ = t.render(cam, s)
        t.upload(sc)
        t.build(args.leaf)
        times = []
        for rep in range(4):
            img, _ = t.render(cam, s)
            f_ms = t.timing()["last_call_ms"]
            w = (2.0 * (img - target)).astype(np.float32)  # dL/dI of L = sum (I - target)^2
            g = t.render_adjoint(cam, s, w)
            a_ms = t.timing()["last_call_ms"]
            if rep:
                times.append((f_ms, a_ms))
        base = sc.materials["color"].copy()
        checks = []
        for m, ch in [(3, 0), (20, 1), (45, 2), (60, 0), (64, 1)]:
            vals = []
            for sign in (1, -1):
                c = base.copy()
                c[m, ch] += np.float32(sign * 1e-3)
                t.set_material_colors(c)
                im, _ = t.render(cam, s)
                vals.append(float(np.sum(w.astype(np.float64) * im)))
            t.set_material_colors(base)
            fd = (vals[0] - vals[1]) / 2e-3
            checks.append({"material": m, "channel": ch, "adjoint": float(g[m, ch]), "fd": fd,
                           "rel_err": abs(float(g[m, ch]) - fd) / max(abs(fd), 1e-12)})
    f_ms = float(np.median([x[0] for x in times]))
    a_ms = float(np.median([x[1] for x in times]))
    cpu = None
    if not args.no_cpu:
        import oracle as O
        threads, host = host_threads()
        t0 = time.time()
        gc = O.render_adjoint(sc, cam, s, w, threads)
        dt = time.time() - t0
        cpu = {"value": round(dt * 1e3, 1), "unit": "ms per adjoint pass", "cores": threads, "host": host, "kind": "port",
               "sample": "the whole adjoint pass (256x256, 32 spp, depth 4)",
               "max_rel_diff_vs_gpu": float(np.max(np.abs(gc - g)) / max(float(np.max(np.abs(gc))), 1e-12)),
               "cosine_similarity_vs_gpu": float(np.sum(gc.astype(np.float64) * g) /
                                                 max(np.linalg.norm(gc.astype(np.float64)) * np.linalg.norm(g.astype(np.float64)), 1e-300))}
    return {"metric": "gradient time, forward + adjoint (64 albedo groups, 256x256, 32 spp, depth 4)",
            "value": round(f_ms + a_ms, 3), "unit": "ms", "forward_ms": round(f_ms, 3), "adjoint_ms": round(a_ms, 3),
            "higher_is_better": False, "dtype": "f32", "data": "synthetic", "n_gpus": 1,
            "fd_checks": checks, "config": {"workload": "Cornell-style box, 64 triangle groups with learnable RGB "
                                                        "albedo, L2 loss against a render with other albedos"},
            "cpu_baseline": cpu}




def _time_mlp(t, cache, sc, P: int):
    """Device time of one inference over P queries and of one 65,536-sample
    training step (forward, backward, Adam), averaged over 20 calls each,
    with events on the context's stream; features uniform in the scene box."""
    import torch
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    t.set_stream(stream.cuda_stream)
    lo, hi = np.asarray(sc.pos.min(axis=0)), np.asarray(sc.pos.max(axis=0))
    g = torch.Generator(device="cuda").manual_seed(0)
    feats = torch.rand((P, 9), device="cuda", generator=g)
    feats[:, :3] = feats[:, :3] * torch.tensor(hi - lo, device="cuda") + torch.tensor(lo, device="cuda")
    out = torch.empty((P, 3), device="cuda")
    tgt = torch.rand((65536, 3), device="cuda", generator=g)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    for _ in range(3):
        cache.infer(feats, out)
    ev[0].record(stream)
    for _ in range(20):
        cache.infer(feats, out)
    ev[1].record(stream)
    tf = feats[:65536].contiguous()
    for _ in range(3):
        cache.train(tf, tgt, want_loss=False)
    ev[2].record(stream)
    for _ in range(20):
        cache.train(tf, tgt, want_loss=False)
    ev[3].record(stream)
    torch.cuda.synchronize()
    t.set_stream(None)
    torch.cuda.set_stream(torch.cuda.default_stream())
    return ev[0].elapsed_time(ev[1]) / 20, ev[2].elapsed_time(ev[3]) / 20


def _time_train_step(t, cache, sc, n: int, reps: int = 10) -> float:
    """Device time of one n-sample training step (forward, backward, Adam),
    averaged over `reps` steps with events on the context's stream."""
    import torch
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    t.set_stream(stream.cuda_stream)
    lo, hi = np.asarray(sc.pos.min(axis=0)), np.asarray(sc.pos.max(axis=0))
    g = torch.Generator(device="cuda").manual_seed(1)
    f = torch.rand((n, 9), device="cuda", generator=g)
    f[:, :3] = f[:, :3] * torch.tensor(hi - lo, device="cuda") + torch.tensor(lo, device="cuda")
    tg = torch.rand((n, 3), device="cuda", generator=g)
    for _ in range(2):
        cache.train(f, tg, want_loss=False)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        cache.train(f, tg, want_loss=False)
    e1.record(stream)
    torch.cuda.synchronize()
    t.set_stream(None)
    torch.cuda.set_stream(torch.cuda.default_stream())
    return e0.elapsed_time(e1) / reps


def bench_mlp(args):
    """The radiance cache's network alone: inference over one 1920x1080
    frame of queries and a 65,536-sample training step, against the BF16
    dense MFMA peak.  Beside the step, a 2^20-sample step: a step's cost is
    a fixed part (the exchange of the workgroups' partial gradients and
    Adam) plus a part per sample, and the large step shows the latter."""
    from master_thesis_amd import scenes
    from master_thesis_amd.cache import RadianceCache
    from master_thesis_amd.tracer import Tracer
    sc = scenes.cornell_closed()
    P = 1920 * 1080
    with Tracer(0) as t:
        t.upload(sc)
        t.build(args.leaf)
        cache = RadianceCache(t, seed=1)
        inf_ms, trn_ms = _time_mlp(t, cache, sc, P)
        big_ms = _time_train_step(t, cache, sc, 1 << 20)
        cache.close()
    big_tf = MLP_TRAIN_FLOP * (1 << 20) / (big_ms * 1e-3) / 1e12
    inf_tf = MLP_FWD_FLOP * P / (inf_ms * 1e-3) / 1e12
    trn_tf = MLP_TRAIN_FLOP * 65536 / (trn_ms * 1e-3) / 1e12
    return {"metric": "TFLOP/s radiance-cache network inference (1920x1080 queries, BF16 MFMA)",
            "value": round(inf_tf, 1), "unit": "TFLOP/s", "higher_is_better": True, "n_gpus": 1,
            "dtype": "bf16 MFMA with f32 accumulation", "data": "synthetic",
            "mlp": {"infer_ms_per_frame": round(inf_ms, 4), "infer_tflops": round(inf_tf, 1),
                    "train_ms_per_step_65536": round(trn_ms, 4), "train_tflops": round(trn_tf, 1)},
            "roofline": {"bound": "mfma", "achieved": round(inf_tf, 1), "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": round(inf_tf / BF16_PEAK_TFLOPS, 4), "traffic": None,
                         "kernel": "k_mlp_infer", "flop_per_sample": MLP_FWD_FLOP},
            "roofline_train": {"bound": "mfma", "achieved": round(trn_tf, 1), "peak": BF16_PEAK_TFLOPS,
                               "unit": "TFLOP/s", "frac": round(trn_tf / BF16_PEAK_TFLOPS, 4),
                               "kernel": "k_mlp_train + k_adam", "flop_per_sample": MLP_TRAIN_FLOP,
                               "step_2p20_samples": {"ms": round(big_ms, 4), "tflops": round(big_tf, 1),
                                                     "frac": round(big_tf / BF16_PEAK_TFLOPS, 4)}},
            "config": {"workload": "78-input MLP 96->64x4->16 (3 outputs), 2,073,600 queries per inference, "
                                   "65,536 samples per training step", "layers": 6}}


# floating-point work of the cache network per sample, on its valid (unpadded)
# inputs and outputs: 78*64 + 4*64*64 + 64*3 = 21,568 multiply-adds forward
MLP_FWD_FLOP = 2 * (78 * 64 + 4 * 64 * 64 + 64 * 3)
# training: forward, weight gradients (the same multiply-adds) and input
# gradients of layers 1..5 (4*64*64 + 64*3)
MLP_TRAIN_FLOP = 2 * MLP_FWD_FLOP + 2 * (4 * 64 * 64 + 64 * 3)
BF16_PEAK_TFLOPS = 2500.0


def bench_denoise(args):
    """Post-filters on a 3840x2160 float RGB image resident in HBM: Gauss
    (kernelSize 5) and the 3x3 median, device time of each kernel from HIP
    events on the context's stream.  Algorithmic bytes: the image read once
    and written once (24 B per pixel).  CPU baseline: the oracle's filters on
    a 480x270 crop, one thread, scaled per pixel."""
    import torch

    from master_thesis_amd.tracer import Tracer
    W, Hh = 3840, 2160
    g = torch.Generator(device="cuda").manual_seed(0)
    img = torch.rand((Hh, W, 3), device="cuda", generator=g)
    res = {}
    with Tracer(0) as t:
        for name, fn in (("gauss5", lambda: t.denoise_gauss(img, 5, 1.5)), ("median3", lambda: t.denoise_median(img))):
            for _ in range(args.warmup):
                fn()
            ms = []
            for _ in range(args.steps):
                fn()
                ms.append(t.timing()["last_kernel_ms"])
            k = float(np.median(ms))
            res[name] = {"kernel_ms": round(k, 4), "gb_per_s": round(24.0 * W * Hh / (k * 1e-3) / 1e9, 1)}
            log(f"[bench] denoise: {name} {k:.4f} ms per 3840x2160 image")
    crop = img[:270, :480].cpu().numpy()
    cpu = {}
    if not args.no_cpu:
        import oracle as O
        for name, fn in (("gauss5", lambda: O.denoise_gauss(crop, 5, 1.5)), ("median3", lambda: O.denoise_median(crop))):
            t0 = time.time()
            fn()
            dt = time.time() - t0
            cpu[name] = round(480 * 270 / dt / 1e6, 2)
    gk = res["gauss5"]["kernel_ms"]
    achieved = res["gauss5"]["gb_per_s"]
    return {"metric": "Mpixels/s, Gauss 5x5 post-filter (3840x2160 float RGB)",
            "value": round(W * Hh / (gk * 1e-3) / 1e6, 1), "unit": "Mpixels/s", "higher_is_better": True,
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "dtype": "f32", "data": "synthetic",
            "filters": res,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "k_gauss",
                         "bytes_alg_per_launch": 24 * W * Hh},
            "cpu_baseline": ({"value": cpu["gauss5"], "unit": "Mpixels/s", "cores": 1, "kind": "port",
                              "sample": "oracle filters on a 480x270 crop of the same image",
                              "median3": cpu["median3"]} if cpu else None),
            "config": {"workload": "GaussDenoiser kernelSize 5 and MedianDenoiser on a 3840x2160 float RGB image"}}


def bench_nrc(args):
    """Learned radiance cache, equal-RMSE protocol.  Scene (default: the
    terrain, sun + sky, depth 8, 1920x1080), a light-sampled reference of
    --ref-spp samples (same expectation as the plain estimator; the plain
    render's mean is reported beside it as the check on its bias), then:
      plain PT          the reference estimator, RMSE at 4 and 16 spp
      light sampling    the cache mode without the network, 1 and 4 spp
      cached            --frames frames of online training at 1 spp (about
                        2^16 training records per frame, 4 Adam steps of
                        16384), then 1-spp frames with training on
      probes            the irradiance-probe field (DDGI) in the cache's
                        place: one field update per frame for --frames
                        frames, then 1-spp frames (update + render), at the
                        cache's query vertex and at the first hit
    Each estimator's RMSE scales as c / sqrt(spp): the plain tracer and light
    sampling are then rendered at the smallest spp that fit predicts for the
    cached frame's RMSE, checked (and raised once if short), and the value is
    the plain tracer's time at that measured spp over the cached frame's time;
    the network's own gain is the same ratio for light sampling.
    The MLP kernels are timed separately on 1920*1080 queries and 65536
    training samples (MFMA throughput against the BF16 dense peak)."""
    import torch

    from master_thesis_amd import scenes, host_lib as H
    from master_thesis_amd._lib import RT_RENDER_SKY, RT_RENDER_SUN
    from master_thesis_amd.cache import RadianceCache
    from master_thesis_amd.tracer import Tracer, make_settings
    W, Hh = (int(x) for x in args.res.split("x"))
    if args.nrc_scene == "terrain":
        sc = scenes.terrain_scene()
        cam = scenes.terrain_cameras(1, W, Hh)[0]
        depth = 8

        def S(spp, off=0, seed=5):
            return make_settings(spp=spp, max_depth=depth, seed=seed, flags=RT_RENDER_SKY | RT_RENDER_SUN,
                                 sample_offset=off, sky=(0.45, 0.55, 0.75), sun_dir=(0.4, 0.8, 0.3),
                                 sun_radius_deg=5.0, sun_radiance=(40.0, 36.0, 30.0))
    else:
        sc = getattr(scenes, args.nrc_scene)()
        closed = args.nrc_scene == "cornell_closed"
        cam = H.closed_camera(W, Hh) if closed else H.cornell_camera(W, Hh)
        depth = 8 if closed else 6

        def S(spp, off=0, seed=5):
            return make_settings(spp=spp, max_depth=depth, seed=seed, sample_offset=off)
    P = W * Hh
    every = max(1, P // 65536)
    res = {}
    with Tracer(0) as t:
        t.upload(sc)
        t.build(args.leaf, cluster=args.nrc_scene == "terrain")
        cache = RadianceCache(t, lr=1e-2, seed=1)
        t0 = time.time()
        ref, _ = cache.render(cam, S(args.ref_spp, 0, 99), train=False, query=False, train_every=0)
        log(f"[bench] nrc: light-sampled reference {args.ref_spp} spp in {time.time() - t0:.1f}s")

        def rmse(img):
            return float(np.sqrt(np.mean((img - ref) ** 2)))

        def run(fn, spp, reps=3):
            e, ms = [], []
            for r in range(reps + 1):
                img, tm = fn(spp, 1000000 + spp * r)
                if r:
                    e.append(rmse(img))
                    ms.append(tm)
            return float(np.mean(e)), float(np.median(ms))

        def plain(spp, off):
            img, _ = t.render(cam, S(spp, off))
            return img, t.timing()["last_call_ms"]

        def light(spp, off):
            img, st = cache.render(cam, S(spp, off), train=False, query=False, train_every=0)
            return img, st.render_ms

        for name, fn, spps in (("plain", plain, (4, 16)), ("light_sampling", light, (1, 4))):
            rows = {}
            for spp in spps:
                e, ms = run(fn, spp)
                rows[spp] = (e, ms)
                log(f"[bench] nrc: {name} {spp} spp rmse {e:.4f} {ms:.2f} ms")
            c = float(np.mean([e * np.sqrt(spp) for spp, (e, _) in rows.items()]))
            ms_spp = rows[spps[-1]][1] / spps[-1]
            res[name] = {"rmse": {str(k): round(v[0], 5) for k, v in rows.items()},
                         "ms": {str(k): round(v[1], 3) for k, v in rows.items()},
                         "c": round(c, 5), "ms_per_spp": round(ms_spp, 4)}
        plain_hi, _ = t.render(cam, S(max(16, args.ref_spp // 4), 0, 98))
        # online training
        cache.close()
        cache = RadianceCache(t, lr=1e-2, seed=1)
        t0 = time.time()
        tms = []
        for f in range(args.frames):
            _, st = cache.render(cam, S(1, f), train_every=every, batch=16384, max_steps=4)
            tms.append(st.render_ms + st.train_ms)
        log(f"[bench] nrc: {args.frames} training frames in {time.time() - t0:.1f}s, last loss {st.loss:.3f}, "
            f"{st.records} records, {st.steps} steps")

        def cached(spp, off):
            img, st = cache.render(cam, S(spp, off), train_every=every, batch=16384, max_steps=4)
            return img, st.render_ms + st.train_ms

        e1, ms1 = run(cached, 1, reps=4)
        log(f"[bench] nrc: cached 1 spp rmse {e1:.4f} {ms1:.2f} ms")
        big, _ = cache.render(cam, S(64, 5000000), train=False, train_every=0)
        res["cached"] = {"rmse": {"1": round(e1, 5)}, "ms": {"1": round(ms1, 3)},
                         "frames_ms_median": round(float(np.median(tms)), 3),
                         "bias": {"mean_diff_vs_reference": round(float((big - ref).mean()), 6),
                                  "rmse_64spp_vs_reference": round(rmse(big), 5),
                                  "light_sampling_expected_rmse_64spp": round(
                                      float(np.hypot(res["light_sampling"]["c"] / 8.0,
                                                     res["light_sampling"]["c"] / np.sqrt(args.ref_spp))), 5)}}
        # plain PT as the check on the reference (and so on the cache's bias): frame means
        res["plain_check"] = {"reference_mean": round(float(ref.mean()), 6), "plain_mean": round(float(plain_hi.mean()), 6),
                        "plain_spp": max(16, args.ref_spp // 4)}
        # irradiance probes (the Vulkan renderer's DDGI term) in the cache's place: a grid over the scene's
        # bounds, one field update per frame (random rotation, multi-bounce, hysteresis), then the same
        # 1-spp frames with the field answering the queries
        from master_thesis_amd.probes import ProbeField
        lo, hi = np.asarray(sc.pos.min(axis=0), np.float64), np.asarray(sc.pos.max(axis=0), np.float64)
        count = (32, 4, 32) if args.nrc_scene == "terrain" else (4, 4, 4)
        inset = 0.05 * (hi - lo)
        spacing = (hi - lo - 2 * inset) / (np.asarray(count) - 1)
        field = ProbeField(t, count, lo + inset, spacing, rays=256, side=8,
                           max_distance=float(np.linalg.norm(hi - lo)), sharpness=50.0,
                           normal_bias=float(0.1 * spacing.min()), hysteresis=0.95, shadow_count=1,
                           multi_bounce=True, rotate=True)
        t0 = time.time()
        ums = []
        for f in range(args.frames):
            ums.append(field.update(S(1, f, 7 + f)))
        log(f"[bench] nrc: {args.frames} probe-field updates in {time.time() - t0:.1f}s "
            f"({float(np.median(ums)):.3f} ms each, {field.num_probes} probes x {field.rays} rays)")
        pr = {"probes": list(count), "rays_per_probe": field.rays, "update_ms_median": round(float(np.median(ums)), 3)}
        for qv in (1, 0):
            def probes(spp, off, qv=qv):
                um = field.update(S(1, 2000000 + off, 3))
                img, st = field.render(cam, S(spp, off), query_vertex=qv)
                return img, st.render_ms + um
            e, ms = run(probes, 1, reps=4)
            bigp, _ = field.render(cam, S(64, 6000000), query_vertex=qv)
            pr[f"query_vertex_{qv}"] = {"rmse": {"1": round(e, 5)}, "ms": {"1": round(ms, 3)},
                                        "rmse_64spp_vs_reference": round(rmse(bigp), 5),
                                        "mean_diff_vs_reference": round(float((bigp - ref).mean()), 6)}
            log(f"[bench] nrc: probes (query vertex {qv}) 1 spp rmse {e:.4f} {ms:.2f} ms, 64 spp rmse {rmse(bigp):.4f}")
        field.close()
        res["probes"] = pr
        # time for each method to reach the cached frame's RMSE, measured: the
        # smallest spp its c / sqrt(spp) fit predicts, rendered and checked
        # against the reference, raised (once, by the measured ratio) if it
        # falls short; the fit's interpolated spp is reported beside it
        ratio, ratio_fit = {}, {}
        for name, fn in (("plain", plain), ("light_sampling", light)):
            need = (res[name]["c"] / e1) ** 2
            n_spp = max(1, int(np.ceil(need)))
            e, ms = run(fn, n_spp, reps=2)
            if e > e1:
                n_spp = max(n_spp + 1, int(np.ceil(n_spp * (e / e1) ** 2)))
                e, ms = run(fn, n_spp, reps=2)
            log(f"[bench] nrc: {name} at {n_spp} spp rmse {e:.4f} (cached {e1:.4f}) {ms:.2f} ms")
            ratio[name] = ms / ms1
            ratio_fit[name] = need * res[name]["ms_per_spp"] / ms1
            res[name]["equal_rmse"] = {"spp": n_spp, "rmse": round(e, 5), "ms": round(ms, 3),
                                       "spp_fit": round(need, 2), "time_ratio_vs_cached": round(ms / ms1, 3),
                                       "time_ratio_fit": round(ratio_fit[name], 3)}
        # MLP kernels on the device, timed with events on the context stream
        inf_ms, trn_ms = _time_mlp(t, cache, sc, P)
        inf_tf = MLP_FWD_FLOP * P / (inf_ms * 1e-3) / 1e12
        trn_tf = MLP_TRAIN_FLOP * 65536 / (trn_ms * 1e-3) / 1e12
        cache.close()
    return {"metric": f"equal-RMSE wall time, plain GPU path tracer / cached form ({args.nrc_scene}, {W}x{Hh})",
            "value": round(ratio["plain"], 2), "unit": "x", "higher_is_better": True, "n_gpus": 1,
            "dtype": "f32 (paths), bf16 MFMA with f32 accumulation (network)", "data": "synthetic",
            # light sampling matches the cached frame's RMSE between 1 and 2 spp, where whole samples
            # overstate its time: its gain is the c / sqrt(spp) fit's, the whole-spp figure beside it
            "network_gain_over_light_sampling": round(ratio_fit["light_sampling"], 3),
            "network_gain_over_light_sampling_whole_spp": round(ratio["light_sampling"], 3),
            "light_sampling_gain_over_plain": round(ratio["plain"] / ratio_fit["light_sampling"], 3),
            "cache_vs_probes": {"rmse_ratio_1spp": round(res["probes"]["query_vertex_1"]["rmse"]["1"] / e1, 3),
                                "ms_ratio_1spp": round(res["probes"]["query_vertex_1"]["ms"]["1"] / ms1, 3)},
            "estimators": res, "training_frames": args.frames, "train_every": every,
            "mlp": {"infer_ms_per_frame": round(inf_ms, 4), "infer_tflops": round(inf_tf, 1),
                    "train_ms_per_step_65536": round(trn_ms, 4), "train_tflops": round(trn_tf, 1)},
            "roofline": {"bound": "mfma", "achieved": round(inf_tf, 1), "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": round(inf_tf / BF16_PEAK_TFLOPS, 4), "traffic": None,
                         "kernel": "k_mlp_infer", "flop_per_sample": MLP_FWD_FLOP},
            "config": {"workload": f"{args.nrc_scene}, {W}x{Hh}, depth {depth}, {args.frames} frames of online "
                                   f"training at 1 spp, reference {args.ref_spp} spp light-sampled",
                       "triangles": int(sc.num_triangles)}}
PMC_PASSES = {  # rocprofv3 passes behind a traversal counter summary, one run each
    "fetch": "--pmc FETCH_SIZE",
    "write": "--pmc WRITE_SIZE TCC_HIT_sum TCC_MISS_sum",
    "sq": "--pmc SQ_INSTS_VALU SQ_ACTIVE_INST_VALU SQ_BUSY_CYCLES SQ_WAVES SQ_WAVE_CYCLES SQ_INSTS_VMEM_RD "
          "SQ_INSTS_SALU SQ_INSTS_LDS TCC_EA0_RDREQ_sum GRBM_GUI_ACTIVE",
    "stats": "--kernel-trace --stats",
}
PMC_WORKLOADS = {"raycast": ("trace_raycast_1M_16M", 16777216), "terrain_raycast": ("trace_terrain_16M", 16588800)}
TRACE_KERNEL = "k_trace<false, false, false>"  # the timed closest-hit walk
STATS_KERNEL = "k_trace<true, false, false>"   # its counting variant, launched once before the timed walks
COPY_KERNEL = "copyBuffer"                      # the 512 MiB calibration copy
CALIB_BYTES = 512 * 1024 * 1024


def pmc_commands(config: str) -> list[str]:
    """The rocprofv3 runs (from /tmp, R = repository root) behind the summary."""
    out = []
    for name, opt in PMC_PASSES.items():
        steps = "--steps 5 --warmup 2" if name == "stats" else "--steps 3 --warmup 1 --calibrate-copy"
        out.append(f"rocprofv3 {opt} --output-format csv -d $R/gpurun_out/pmc -o {config}_{name} -- "
                   f"python3 $R/bench.py --config {config} {steps} --no-cpu --no-halves")
    return out


def summarize_pmc(config: str, pmc_dir: Path) -> None:
    """Turns the rocprofv3 CSV files of pmc_commands(config) under pmc_dir into
    profiles/<name>_pmc.json, _pmc.md and _kernel_stats.md: per launch of the
    timed closest-hit walk, the memory-side bytes (FETCH_SIZE and WRITE_SIZE
    scaled by what they read for a 512 MiB device copy in the same run), the
    L2 hit rate, instruction counts and the VALU issue rate."""
    import csv
    name, rays = PMC_WORKLOADS[config]

    def rows(kind: str, suffix: str):
        files = sorted(pmc_dir.rglob(f"{config}_{kind}*{suffix}"))
        if not files:
            raise SystemExit(f"no {config}_{kind}*{suffix} under {pmc_dir}")
        for f in files:
            with open(f, newline="") as fh:
                yield from csv.DictReader(fh)

    def after_counting_pass(rs) -> int:
        # the timed launches follow the counting walk's one launch; walks before
        # it (the terrain's own ray generation) trace other rays on another tree
        ids = [int(r["Dispatch_Id"]) for r in rs if STATS_KERNEL in r["Kernel_Name"]]
        if not ids:
            raise SystemExit("no counting launch in the profile")
        return max(ids)

    def counters(kind: str) -> dict:
        # per dispatch: {counter: value} for the walk's and the copy's launches
        per, rs = {}, list(rows(kind, "counter_collection.csv"))
        first = after_counting_pass(rs)
        for r in rs:
            k = r["Kernel_Name"]
            tag = "trace" if TRACE_KERNEL in k else ("copy" if COPY_KERNEL in k else None)
            if tag is None or (tag == "trace" and int(r["Dispatch_Id"]) <= first):
                continue
            d = per.setdefault((tag, r["Dispatch_Id"]), {})
            d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
        trace = [v for (tag, _), v in per.items() if tag == "trace"]
        copy = [v for (tag, _), v in per.items() if tag == "copy" and v.get("FETCH_SIZE", v.get("WRITE_SIZE", 0)) > 1e5]
        if not trace:
            raise SystemExit(f"no {TRACE_KERNEL} launches in the {kind} pass")
        mean = lambda vs, c: float(np.mean([v[c] for v in vs]))  # noqa: E731
        return {"trace": trace, "copy": copy, "mean": mean}

    fetch, write, sq = counters("fetch"), counters("write"), counters("sq")
    m = fetch["mean"]
    fetch_kib = m(fetch["trace"], "FETCH_SIZE")
    fetch_factor = CALIB_BYTES / 1024 / m(fetch["copy"], "FETCH_SIZE")
    write_kib = m(write["trace"], "WRITE_SIZE")
    write_factor = CALIB_BYTES / 1024 / m(write["copy"], "WRITE_SIZE")
    hbm = (fetch_kib * fetch_factor + write_kib * write_factor) * 1024
    hit, miss = m(write["trace"], "TCC_HIT_sum"), m(write["trace"], "TCC_MISS_sum")
    s = {c: m(sq["trace"], c) for c in ("SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_VMEM_RD", "SQ_INSTS_LDS",
                                          "TCC_EA0_RDREQ_sum", "GRBM_GUI_ACTIVE", "SQ_BUSY_CYCLES",
                                          "SQ_ACTIVE_INST_VALU", "SQ_WAVES", "SQ_WAVE_CYCLES")}
    cycles = s["GRBM_GUI_ACTIVE"] / 8  # GRBM_GUI_ACTIVE sums the 8 XCDs
    valu_issue = s["SQ_INSTS_VALU"] / (256 * cycles)
    valu_busy = s["SQ_ACTIVE_INST_VALU"] / (256 * cycles)

    stats = list(rows("stats", "kernel_stats.csv"))
    kt = list(rows("stats", "kernel_trace.csv"))
    first = after_counting_pass(kt)
    launches = [r for r in kt if TRACE_KERNEL in r["Kernel_Name"] and int(r["Dispatch_Id"]) > first]
    durs = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6 for r in launches]
    k_ms = float(np.mean(durs))

    out = {"kernel": TRACE_KERNEL, "workload": config, "rays_per_launch": rays, "source_hash": kernel_source_hash(),
           "launches_counted": len(fetch["trace"]), "fetch_kib_raw": round(fetch_kib, 1),
           "write_kib_raw": round(write_kib, 1), "fetch_factor": round(fetch_factor, 4),
           "write_factor": round(write_factor, 4), "hbm_bytes_per_launch": int(hbm),
           "bytes_per_ray": round(hbm / rays, 1), "tcc_hit": int(hit), "tcc_miss": int(miss),
           "l2_hit_rate": round(hit / (hit + miss), 4), "tcc_ea0_rdreq": int(s["TCC_EA0_RDREQ_sum"]),
           "sq_insts_valu": int(s["SQ_INSTS_VALU"]), "sq_insts_salu": int(s["SQ_INSTS_SALU"]),
           "sq_insts_vmem_rd": int(s["SQ_INSTS_VMEM_RD"]), "sq_insts_lds": int(s["SQ_INSTS_LDS"]),
           "grbm_gui_active_per_xcd": int(cycles), "cus": 256, "xcds": 8,
           "valu_insts_per_cu_cycle": round(valu_issue, 4), "valu_busy": round(valu_busy, 4),
           "kernel_ms_trace": round(k_ms, 4),
           "kernel_launches_traced": len(durs)}
    prof = ROOT / "profiles"
    (prof / f"{name}_pmc.json").write_text(json.dumps(out, indent=1) + "\n")
    cmds = "\n".join(pmc_commands(config))
    (prof / f"{name}_pmc.md").write_text(
        f"# Counters: {TRACE_KERNEL} on {config} ({rays} rays per launch)\n\n"
        f"Kernel sources hashed {out['source_hash']} (`bench.py` `KERNEL_SOURCES`); the bench uses these figures "
        f"only for the same hash.\n\nCommands (one pass each, `R` = repository root, run from `/tmp`), then "
        f"`python3 bench.py --config {config} --summarize-pmc $R/gpurun_out/pmc`:\n\n```\n{cmds}\n```\n\n"
        f"Timed launches of the walk (the counting pass's launch is another kernel), mean per launch:\n\n"
        f"| figure | value |\n|---|---|\n"
        f"| kernel duration (kernel trace, {len(durs)} launches) | {k_ms:.4f} ms |\n"
        f"| FETCH_SIZE raw / calibrated | {fetch_kib:.0f} KiB × {fetch_factor:.3f} |\n"
        f"| WRITE_SIZE raw / calibrated | {write_kib:.0f} KiB × {write_factor:.3f} |\n"
        f"| memory-side bytes per launch | {hbm / 1e9:.2f} GB ({hbm / rays:.0f} B per ray) |\n"
        f"| memory-side rate at the traced duration | {hbm / k_ms / 1e9:.2f} TB/s ({hbm / k_ms / 1e9 / 8:.2f} of 8 TB/s) |\n"
        f"| L2 hits / misses (TCC_HIT_sum / TCC_MISS_sum) | {hit:.0f} / {miss:.0f}: hit rate {hit / (hit + miss):.3f} |\n"
        f"| VALU instructions (SQ_INSTS_VALU, wave64) | {s['SQ_INSTS_VALU']:.0f} ({s['SQ_INSTS_VALU'] * 64 / rays:.0f} lane-slots per ray) |\n"
        f"| SALU / VMEM read / LDS instructions | {s['SQ_INSTS_SALU']:.0f} / {s['SQ_INSTS_VMEM_RD']:.0f} / {s['SQ_INSTS_LDS']:.0f} |\n"
        f"| cycles per XCD (GRBM_GUI_ACTIVE / 8) | {cycles:.0f} |\n"
        f"| VALU issue rate: SQ_INSTS_VALU / (256 CUs × cycles) | {valu_issue:.3f} wave64 instructions per CU per cycle |\n"
        f"| VALUBusy (rocprof's gfx950 form: SQ_ACTIVE_INST_VALU / 256 CUs / cycles) | {valu_busy:.3f} |\n\n"
        f"A wave64 VALU instruction that holds its SIMD for 4 cycles allows one instruction per CU per cycle; the "
        f"measured rate reaching or passing 1 (some of gfx950's VALU operations complete in fewer cycles) says the "
        f"VALU pipes are saturated.\n"
        f"Calibration: a 512 MiB torch device copy (`__amd_rocclr_copyBuffer`) in the same runs reads FETCH_SIZE "
        f"{m(fetch['copy'], 'FETCH_SIZE'):.0f} KiB (factor {fetch_factor:.3f}) and WRITE_SIZE "
        f"{m(write['copy'], 'WRITE_SIZE'):.0f} KiB (factor {write_factor:.3f}).\n"
        f"Memory-side bytes are L2 misses to the fabric: the 256 MB MALL serves part of them, so HBM moved at most "
        f"this much.\n")
    tot = sum(float(r["TotalDurationNs"]) for r in stats)
    table = "\n".join(f"| `{_short_kernel(r['Name'])}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.3f} | "
                      f"{float(r['AverageNs']) / 1e6:.4f} | {100 * float(r['TotalDurationNs']) / tot:.2f} |"
                      for r in sorted(stats, key=lambda r: -float(r["TotalDurationNs"]))[:16])
    (prof / f"{name}_kernel_stats.md").write_text(
        f"# Kernel stats: bench.py --config {config} (rocprofv3 --kernel-trace --stats)\n\n"
        f"```\ncd /tmp && {pmc_commands(config)[-1]}\n```\n\n"
        f"Launches of {TRACE_KERNEL} over all {rays} rays (after the counting pass): {len(durs)}, mean "
        f"{k_ms:.4f} ms = {rays / k_ms / 1e3:.1f} Mrays/s.\n\n"
        f"| kernel | calls | total ms | mean ms | % |\n|---|---|---|---|---|\n{table}\n")
    log(f"[pmc] {name}: {json.dumps(out)}")


def _short_kernel(k: str) -> str:
    return k.replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "").replace("rt::", "")


def summarize_stats(d: Path, name: str, cmd: str) -> None:
    """profiles/<name>_kernel_stats.md: every kernel's calls and time from the
    *_kernel_stats.csv files rocprofv3 wrote under d."""
    import csv
    rows = []
    for f in sorted(d.rglob("*kernel_stats.csv")):
        with open(f, newline="") as fh:
            rows += list(csv.DictReader(fh))
    if not rows:
        raise SystemExit(f"no kernel_stats.csv under {d}")
    tot = sum(float(r["TotalDurationNs"]) for r in rows)
    table = "\n".join(f"| `{_short_kernel(r['Name'])}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.3f} | "
                      f"{float(r['AverageNs']) / 1e3:.1f} | {100 * float(r['TotalDurationNs']) / tot:.1f} |"
                      for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])))
    (ROOT / "profiles" / f"{name}_kernel_stats.md").write_text(
        f"# Kernel trace: {name}\n\nCommand (run from /tmp on the MI355X box, repo at $R), then "
        f"`python3 bench.py --summarize-stats DIR {name} CMD`:\n\n```\ncd /tmp && {cmd}\n```\n\n"
        f"| kernel | calls | total ms | avg us | % |\n|---|---|---|---|---|\n{table}\n")


if __name__ == "__main__":
    sys.exit(main())
