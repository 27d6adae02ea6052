"""Headline benchmark: Mrays/s of the path-tracing hot path on one MI355X.

Workload (one "step" = one frame): 1920x1080, 64 spp, maxPathLength 8 with
Russian roulette from bounce 3, over 1,004,098 triangles (2000 seeded displaced
spheres of 500 triangles + ground + 4096 emissive triangles in 256 emitters)
under a procedural 1024x512 HDR sky sampled with an alias table.  Synthetic
seeded scene, inputs resident in HBM before the timed region.

Prints one JSON line with the driver's contract plus:
  roofline      the dominant kernel's algorithmic HBM bytes per launch / its
                mean HIP-event duration, against 8 TB/s
  cpu_baseline  the CPU oracle (a C port of the reference path tracer) on this
                host's cores over a bounded crop of the same frame
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--samples-per-wave", type=int, default=0, help="0: library default")
    ap.add_argument("--no-restir", action="store_true", help="plain NEE instead of ReSTIR DI")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the renderer shards by image tiles or sample ranges, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from stratum2_amd import abi, scenes
    from stratum2_amd.renderer import Context, Scene, TestRenderer

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    t0 = time.time()
    sd = scenes.spheres_field(seed=1)
    view = scenes.spheres_view(args.width, args.height)
    log(f"scene generated: {sd.triangle_count} triangles in {time.time() - t0:.1f}s")
    ctx = Context(0)
    t0 = time.time()
    sc = Scene(ctx, sd)
    log(f"scene uploaded + BVH built in {time.time() - t0:.1f}s: {sc.stats()}")
    r = TestRenderer(sc)
    flags = abi.ST_FLAG_DEFAULT | abi.ST_FLAG_ENV_ALIAS
    if not args.no_restir:  # ReSTIR DI: 32 candidates, temporal + spatial (5 neighbours, 30 px) reuse
        flags |= abi.ST_FLAG_RESTIR_DI | abi.ST_FLAG_RESTIR_DI_REUSE
    params = abi.default_params(args.width, args.height, max_depth=8, spp=args.spp, flags=flags, rr_start_depth=3,
                                samples_per_wave=args.samples_per_wave)

    # traversal statistics for the algorithmic-bytes model (untimed, 1 spp)
    cp = abi.default_params(args.width, args.height, max_depth=8, spp=1, flags=flags | abi.ST_FLAG_COUNT_TRAVERSAL,
                            rr_start_depth=3)
    r.render(view, cp)
    torch.cuda.synchronize()
    tc = ctx.timings()
    nodes_c = tc["closest_nodes"] / max(1, tc["closest_rays"])
    tris_c = tc["closest_tris"] / max(1, tc["closest_rays"])
    nodes_s = tc["shadow_nodes"] / max(1, tc["shadow_rays"])
    tris_s = tc["shadow_tris"] / max(1, tc["shadow_rays"])
    log(f"per closest ray: {nodes_c:.1f} nodes {tris_c:.1f} tris; per shadow ray: {nodes_s:.1f} nodes {tris_s:.1f} tris")

    out = None
    for _ in range(args.warmup):
        out = r.render(view, params, out=out)
    torch.cuda.synchronize()
    times, rays, kt = [], [], []
    for _ in range(args.steps):
        torch.cuda.synchronize()
        a = time.perf_counter()
        out = r.render(view, params, out=out)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - a)
        rc = out["ray_counts"].cpu().numpy().view(np.uint64)
        rays.append(int(rc[0]) + int(rc[1]))
        kt.append(ctx.timings())
        log(f"step: {times[-1] * 1e3:.1f} ms, {rays[-1] / 1e6:.1f} M rays, kernels {kt[-1]}")
    total_s = sum(times)
    value = sum(rays) / total_s / 1e6
    ms = total_s / len(times) * 1e3

    # dominant kernel class and its algorithmic bytes: 32 B ray + 16 B hit (or
    # 48 B shadow record + 4 B flag) + 48 B per 4-wide node + 48 B per triangle
    NODE_B, TRI_B = 48, 48
    t_last = kt[-1]
    classes = {
        "trace_closest": (t_last["trace_ms"], t_last["trace_launches"], t_last["closest_rays"] * (48 + NODE_B * nodes_c + TRI_B * tris_c)),
        "trace_shadow": (t_last["shadow_ms"], t_last["shadow_launches"], t_last["shadow_rays"] * (52 + NODE_B * nodes_s + TRI_B * tris_s)),
        "shade": (t_last["shade_ms"], t_last["shade_launches"], t_last["closest_rays"] * (16 + 64 + 64 + 64 + 48) + t_last["shadow_rays"] * 48),
    }
    dom = max(classes, key=lambda k: classes[k][0])
    dms, dl, dbytes = classes[dom]
    achieved = dbytes / (dms / 1e3) / 1e9 if dms > 0 else 0.0
    roofline = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "bytes_per_launch": round(dbytes / max(1, dl)), "avg_launch_ms": round(dms / max(1, dl), 4),
                "per_ray": {"closest_nodes": round(nodes_c, 2), "closest_tris": round(tris_c, 2),
                            "shadow_nodes": round(nodes_s, 2), "shadow_tris": round(tris_s, 2)}}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(sd, view, args, log)

    line = {
        "metric": "Mrays/s 1080p/1M tris", "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded scene generator)",
        "config": {"workload": ("1920x1080 64spp maxPathLength 8 (RR from 3), 1,004,098 tris, 4096 emissive tris, "
                                "1024x512 env alias-sampled" if (args.width, args.height, args.spp) == (1920, 1080, 64)
                                else f"{args.width}x{args.height} {args.spp}spp maxPathLength 8")
                               + ("" if args.no_restir else ", ReSTIR DI 32 candidates + temporal + spatial 5x30px"),
                   "restir_di": not args.no_restir,
                   "width": args.width, "height": args.height, "spp": args.spp, "max_path_length": 8,
                   "triangles": sd.triangle_count, "samples_per_px_per_s": round(args.width * args.height * args.spp / (ms / 1e3), 1)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def cpu_baseline(sd, view, args, log):
    """Oracle (C port of the reference tracer, OpenMP) on the same frame at a
    bounded number of samples (reference semantics: mip-warp env, RR at every
    bounce), grown until it takes about a third of --cpu-seconds."""
    from stratum2_amd import abi
    from oracle.oracle import OracleScene
    o = OracleScene(sd)
    threads = min(16, os.cpu_count() or 1)  # this host's share of cores
    spp = 1
    while True:
        p = abi.default_params(args.width, args.height, max_depth=8, spp=spp)
        a = time.perf_counter()
        res = o.render(view, p, threads=threads, want_path_length=False)
        dt = time.perf_counter() - a
        nr = int(res["ray_counts"][0]) + int(res["ray_counts"][1])
        if dt >= args.cpu_seconds / 3 or spp >= 64:
            break
        spp = min(64, spp * max(2, int(args.cpu_seconds / 3 / max(dt, 1e-3)) + 1))
    log(f"cpu baseline: {args.width}x{args.height} x {spp} spp, {nr} rays in {dt:.2f}s on {threads} threads")
    return {"value": round(nr / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"CPU oracle (C port of the reference tracer), same {args.width}x{args.height} frame at {spp} spp, "
                      f"maxPathLength 8, {nr} rays in {dt:.1f} s"}


if __name__ == "__main__":
    main()
