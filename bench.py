"""Headline benchmark: Mrays/s of the path-tracing hot path on one MI355X.

Workload (one "step" = one frame), --workload spheres (default): 1920x1080,
64 spp, maxPathLength 8 with Russian roulette from bounce 3, over 1,004,098
triangles (2000 seeded displaced spheres of 500 triangles + ground + 4096
emissive triangles in 256 emitters) under a procedural 1024x512 HDR sky
sampled with an alias table, ReSTIR DI (32 candidates, temporal + spatial
reuse over 5 neighbours in 30 px).  --workload cloud: 1024x1024, 32 spp,
maxPathLength 16, a 256^3 NanoVDB fBm density grid (sigma_t scale 20, albedo
0.9, g 0.3) under the same sky, delta tracking + ratio tracking.  Synthetic
seeded scenes, inputs resident in HBM before the timed region.

Prints one JSON line with the driver's contract plus:
  roofline      the dominant kernel's algorithmic HBM bytes per launch / its
                mean HIP-event duration, against 8 TB/s; traffic = the HBM
                bytes per launch measured for that kernel by rocprofv3 PMC
                (FETCH_SIZE, WRITE_SIZE) and recorded in profiles/
  cpu_baseline  the CPU oracle (a C port of the reference path tracer) on this
                host's cores over a bounded sample of the same frame
--light-trace times TestRenderer's light-tracing mode on the spheres workload
instead (light subpaths, one per pixel and sample, connected to the camera);
--bdpt its bidirectional mode with a light vertex cache (gUseVC).
--workload svgf times the SVGF denoiser (st_denoise) on a 1920x1080 frame of
the spheres workload: temporal accumulation with reprojection and albedo
demodulation, variance estimate, five 5x5 a-trous passes (one step = one
denoised frame, Mpixels/s).
--workload batch times one st_batch_render of 32 seeded requests of uneven
size and spp (square, 128 to 1024 px a side; 1, 4, 16 or 64 spp): three in
four render the spheres scene as the headline does, one in four the cloud as
--workload cloud does (one step = one batch, Mrays/s over all requests),
beside the same requests rendered one after another with st_render.  Launch
shapes come from the package's autotuner, run before the timed region
(--no-autotune: the library defaults).
--workload adjoint times the path-replay adjoint (st_adjoint): one step = the
gradient of an image loss over 16 cameras at 512x512, 8 spp, maxPathLength 8
(RR from 3), with respect to the per-triangle albedo (5144 triangles) and a
64^3 density grid of the Cornell box with a cloud; reported in ms per
gradient and as a multiple of the 16 forward renders.
--workload cache measures time-to-RMSE of the learned radiance cache
(st_render_cached): paths end in the cache at their second vertex, 8-spp
frames each train one Adam step on 2^16 rows; reported as the time the cached
form needs to reach the RMSE the plain tracer has after 128 spp, over the
plain tracer's time (a 1024-spp plain render is the reference), on the volume
scene (the value) and the 1M-triangle headline scene.
--copy-calibration runs a plain device copy of known size (for scaling the
PMC byte counters) and exits.
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
    ap.add_argument("--restir-deep", action="store_true",
                    help="RIS at the NEE vertices past the primary one too (ST_FLAG_RESTIR_DI_DEEP)")
    ap.add_argument("--workload", choices=["spheres", "cloud", "svgf", "batch", "adjoint", "cache"], default="spheres")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="batch workload: the library's default launch shapes")
    ap.add_argument("--device-bvh", action="store_true", help="rebuild the BVH on the GPU (st_scene_update) before timing")
    ap.add_argument("--light-trace", action="store_true", help="light tracing (paths start on the lights)")
    ap.add_argument("--bdpt", action="store_true", help="bidirectional mode with a light vertex cache (gUseVC)")
    ap.add_argument("--serial", action="store_true",
                    help="time the frames with waves one after another on one stream (ST_FLAG_SERIAL_WAVES), "
                         "so a profiler sees each kernel alone")
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

    if args.copy_calibration:
        copy_calibration(log)
        return
    if args.workload == "svgf":
        bench_svgf(args, log)
        return
    if args.workload == "batch":
        bench_batch(args, log)
        return
    if args.workload == "cache":
        bench_cache(args, log)
        return
    if args.workload == "adjoint":
        bench_adjoint(args, log)
        return
    cloud = args.workload == "cloud"
    if cloud:
        if args.width == 1920 and args.height == 1080:
            args.width = args.height = 1024
        if args.spp == 64:
            args.spp = 32
        args.no_restir = True  # ReSTIR DI is not offered with media
    if args.light_trace or args.bdpt:
        if cloud:
            print("--light-trace and --bdpt run the spheres workload (neither is offered with media)", file=sys.stderr)
            sys.exit(2)
        args.no_restir = True  # the reference turns ReSTIR DI off when light tracing; the bidirectional mode uses plain NEE
    max_depth = 16 if cloud else 8

    t0 = time.time()
    if cloud:
        sd = scenes.cloud(seed=0)
        view = scenes.cloud_view(args.width, args.height)
    else:
        sd = scenes.spheres_field(seed=1)
        view = scenes.spheres_view(args.width, args.height)
    log(f"scene generated: {sd.triangle_count} triangles, {len(sd.volumes)} volumes in {time.time() - t0:.1f}s")
    ctx = Context(0)
    t0 = time.time()
    sc = Scene(ctx, sd)
    log(f"scene uploaded + BVH built in {time.time() - t0:.1f}s: {sc.stats()}")
    bvh_ms = None
    if args.device_bvh:
        import numpy as _np
        bvh_ms = sc.update(_np.stack(sd.transforms))
        log(f"device BVH rebuild: {bvh_ms:.2f} ms: {sc.stats()}")
    r = TestRenderer(sc)
    flags = abi.ST_FLAG_DEFAULT | abi.ST_FLAG_ENV_ALIAS
    if args.light_trace:
        flags |= abi.ST_FLAG_TRACE_FROM_LIGHT
    if args.bdpt:
        flags |= abi.ST_FLAG_BDPT_LVC
    if not args.no_restir:  # ReSTIR DI: 32 candidates, temporal + spatial (5 neighbours, 30 px) reuse
        flags |= abi.ST_FLAG_RESTIR_DI | abi.ST_FLAG_RESTIR_DI_REUSE
        if args.restir_deep:
            flags |= abi.ST_FLAG_RESTIR_DI_DEEP
    # roulette from path length 3 (north star); every vertex for the volume
    # and for the bidirectional mode (whose MIS weights assume it)
    rr = 0 if (cloud or args.bdpt) else 3
    params = abi.default_params(args.width, args.height, max_depth=max_depth, spp=args.spp,
                                flags=flags | (abi.ST_FLAG_SERIAL_WAVES if args.serial else 0), rr_start_depth=rr,
                                samples_per_wave=args.samples_per_wave)

    # traversal statistics for the algorithmic-bytes model (untimed, 1 spp);
    # with media the node counters count density-grid reads
    cp = abi.default_params(args.width, args.height, max_depth=max_depth, spp=1, flags=flags | abi.ST_FLAG_COUNT_TRAVERSAL,
                            rr_start_depth=rr)
    first = r.render(view, cp)
    torch.cuda.synchronize()
    tc = ctx.timings()
    # primary hits per sample (unjittered camera: the same for every sample)
    prim_hits = int((first["visibility"][..., 0] != -1).sum().item())
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

    # Exclusive per-kernel times: one more frame, untimed, with the waves one
    # after another on one stream (the same image), HIP events around each
    # launch.  The dominant kernel is the one with the most device time.
    serial = abi.default_params(args.width, args.height, max_depth=max_depth, spp=args.spp,
                                flags=flags | abi.ST_FLAG_SERIAL_WAVES, rr_start_depth=rr,
                                samples_per_wave=args.samples_per_wave)
    a = time.perf_counter()
    r.render(view, serial, out=out)
    torch.cuda.synchronize()
    serial_ms = (time.perf_counter() - a) * 1e3
    kx = ctx.kernel_times()
    ts = ctx.timings()
    log(f"serial frame {serial_ms:.1f} ms, per kernel (ms, launches): {kx}")

    algo, names = algo_bytes(ts, (nodes_c, tris_c, nodes_s, tris_s), cloud, prim_hits, args.width * args.height * args.spp,
                             params, not args.no_restir, args.light_trace, args.bdpt)
    kernels = {}
    for k, b in algo.items():
        ms_k, n_k = kx[k]
        if n_k == 0 or ms_k <= 0:
            continue
        kernels[names.get(k, k)] = {"ms_per_frame": round(ms_k, 3), "launches": n_k, "avg_launch_ms": round(ms_k / n_k, 4),
                                    "bytes_per_launch": round(b / n_k), "frac": round(b / (ms_k / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_frame"])
    d = kernels[dom]
    # exclusive launches: the kernels' own times add up to at most the serial
    # frame.  (They may exceed the timed frame, where four waves overlap and
    # hide each other's latency.)
    assert sum(k["ms_per_frame"] for k in kernels.values()) <= serial_ms, \
        f"{sum(k['ms_per_frame'] for k in kernels.values()):.1f} ms of kernel time in a {serial_ms:.1f} ms serial frame"
    achieved = d["bytes_per_launch"] / (d["avg_launch_ms"] / 1e3) / 1e9
    traffic, traffic_src = measured_traffic(args, dom)
    roofline = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_src,
                "bytes_per_launch": d["bytes_per_launch"], "avg_launch_ms": d["avg_launch_ms"],
                "timing": "HIP events around each launch in a frame rendered with ST_FLAG_SERIAL_WAVES "
                          f"(no kernel overlap; {serial_ms:.1f} ms for the whole frame)",
                "kernels": kernels,
                "per_ray": ({"closest_grid_reads": round(nodes_c, 2), "shadow_grid_reads": round(nodes_s, 2)} if cloud else
                            {"closest_nodes": round(nodes_c, 2), "closest_tris": round(tris_c, 2),
                             "shadow_nodes": round(nodes_s, 2), "shadow_tris": round(tris_s, 2)})}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(sd, view, args, max_depth, flags, rr, log)

    if cloud:
        workload = (f"{args.width}x{args.height} {args.spp}spp maxPathLength 16, 256^3 NanoVDB fBm grid (sigma_t 20, "
                    "albedo 0.9, g 0.3), 1024x512 env alias-sampled, delta + ratio tracking")
    elif (args.width, args.height, args.spp) == (1920, 1080, 64):
        workload = ("1920x1080 64spp maxPathLength 8 (RR from 3), 1,004,098 tris, 4096 emissive tris, "
                    "1024x512 env alias-sampled")
    else:
        workload = f"{args.width}x{args.height} {args.spp}spp maxPathLength 8"
    if not cloud and not args.no_restir:
        workload += (", ReSTIR DI at primary vertices (32 candidates, temporal + spatial reuse over 5 neighbours in 30 px), " +
                     ("RIS over 32 candidates at deeper vertices" if args.restir_deep else "one-sample NEE deeper"))
    if args.light_trace:
        workload += ", light tracing (1 light subpath per pixel and sample, connected to the camera)"
    if args.bdpt:
        workload += ", bidirectional with a light vertex cache (1 light subpath per pixel and sample)"
    line = {
        "metric": "Mrays/s volume 1024^2" if cloud else "Mrays/s 1080p/1M tris", "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded scene generator)",
        "config": {"workload": workload, "restir_di": not args.no_restir, "restir_deep": args.restir_deep and not args.no_restir,
                   "light_tracing": args.light_trace,
                   "bdpt": args.bdpt,
                   "width": args.width, "height": args.height, "spp": args.spp, "max_path_length": max_depth,
                   "triangles": sd.triangle_count, "bvh": "device linear BVH" if args.device_bvh else "host binned SAH",
                   "bvh_build_ms": round(bvh_ms, 2) if bvh_ms is not None else None, "samples_per_px_per_s": round(args.width * args.height * args.spp / (ms / 1e3), 1)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def algo_bytes(ts, per_ray, cloud, prim_hits, npix_spp, params, restir, light_trace=False, bdpt=False):
    """Algorithmic HBM bytes of one frame per kernel class, from its ray counts
    (ts: Context.timings()) and the per-ray traversal counts of a
    COUNT_TRAVERSAL pass (per_ray: closest nodes, closest triangles, shadow
    nodes, shadow triangles; density-grid reads with media).

    Traversal: 32 B ray + 16 B hit (or 48 B shadow record + 4 B flag) + 48 B
    per 4-wide node + 48 B per triangle.  Media (per traceRay segment): 52 B
    path state read + 52 B written (or the 48 B shadow record + 16 B NEE
    accumulator), and 44 B per density-grid read (root tile key + child, upper
    mask word + table entry, lower mask word + table entry, leaf value).
    Shading: 16 B hit + 64 B path state + 64 B material + 64 B triangle + 48 B
    shadow record per vertex.  k_ris, per primary hit: the same 208 B of hit,
    state, material and triangle, `cand` light candidates (environment, alias
    mode: 8 B alias entry + 4 B pdf + 4 texels of 16 B; emissive triangle: 16 B
    light record + 80 B emitter triangle + 80 B transform and material words),
    mixed by the environment sampling probability, and its 48 B reservoir +
    48 B record written.  Reuse, per pixel and sample: temporal 144 B (two
    reservoirs, the primary record, one written), spatial 160 B (own reservoir
    and record read, one reservoir and one shadow record written); the
    neighbours' records are other pixels' reads, and light-sample
    re-evaluations read the same few MB of emitter table and environment as
    k_ris, so neither is counted again.  Returns ({class: bytes}, {class:
    kernel name where it differs})."""
    NODE_B, TRI_B, GRID_B = 48, 48, 44
    CAND_ENV_B, CAND_TRI_B = 76, 176
    nodes_c, tris_c, nodes_s, tris_s = per_ray
    env_p = params.environment_sample_probability
    cand_b = env_p * CAND_ENV_B + (1 - env_p) * CAND_TRI_B
    cr, sr = ts["closest_rays"], ts["shadow_rays"]
    if cloud:
        algo = {"k_trace_closest": cr * 104 + GRID_B * nodes_c * cr,
                "k_trace_shadow": sr * 64 + GRID_B * nodes_s * sr,
                "k_shade": cr * (16 + 64 + 64 + 28 + 48) + sr * 48}
        names = {"k_trace_closest": "k_trace_media", "k_trace_shadow": "k_shadow_media"}
    else:
        algo = {"k_trace_closest": cr * (48 + NODE_B * nodes_c + TRI_B * tris_c),
                "k_trace_shadow": sr * (52 + NODE_B * nodes_s + TRI_B * tris_s),
                "k_shade": cr * (16 + 64 + 64 + 64 + 48) + sr * 48}
        names = {"k_shade": "k_shade_light" if light_trace else "k_shade_vc" if bdpt else "k_shade"}
        if bdpt:  # a cached-vertex connection adds its 48 B vertex, 64 B triangle and 64 B material
            algo["k_shade"] += sr * (48 + 64 + 64) / 2
        if restir:
            algo["k_ris"] = prim_hits * (npix_spp // (params.width * params.height)) * \
                (208 + params.candidate_samples * cand_b + 96)
            algo["k_restir_reuse"] = npix_spp * (144 + 160)
    return algo, names


def batch_requests(n=32, seed=7):
    """The batch mix: n (scene kind, side, spp) triples drawn from a seeded
    generator, every fourth request on the cloud."""
    import numpy as np
    rng = np.random.default_rng(seed)
    sides = rng.choice([128, 256, 512, 1024], n)
    spps = rng.choice([1, 4, 16, 64], n)
    return [("cloud" if i % 4 == 3 else "spheres", int(sides[i]), int(spps[i])) for i in range(n)]


def bench_batch(args, log):
    """One step = one st_batch_render of the batch mix (one hipGraph launch
    spread over four streams).  Mrays/s counts every request's rays.  The same
    requests rendered one after another with st_render are timed beside it;
    per-kernel roofline figures come from each request rendered once more with
    ST_FLAG_SERIAL_WAVES (exclusive HIP-event times, summed over requests)."""
    import numpy as np
    import torch
    from stratum2_amd import abi, scenes
    from stratum2_amd.renderer import Batch, Context, Scene, TestRenderer
    mix = batch_requests()
    ctx = Context(0)
    t0 = time.time()
    sd = {"spheres": scenes.spheres_field(seed=1), "cloud": scenes.cloud(seed=0)}
    sc = {k: Scene(ctx, v) for k, v in sd.items()}
    log(f"scenes built in {time.time() - t0:.1f}s")
    head = abi.ST_FLAG_DEFAULT | abi.ST_FLAG_ENV_ALIAS | abi.ST_FLAG_RESTIR_DI | abi.ST_FLAG_RESTIR_DI_REUSE
    reqs = []
    for kind, side, spp in mix:
        if kind == "cloud":
            p = abi.default_params(side, side, max_depth=16, spp=spp, flags=abi.ST_FLAG_DEFAULT | abi.ST_FLAG_ENV_ALIAS,
                                   rr_start_depth=0)
            reqs.append((sc[kind], scenes.cloud_view(side, side), p))
        else:
            p = abi.default_params(side, side, max_depth=8, spp=spp, flags=head, rr_start_depth=3)
            reqs.append((sc[kind], scenes.spheres_view(side, side), p))
    tune = None
    if not args.no_autotune:  # launch shapes for this mix (untimed; fresh batches)
        from stratum2_amd.autotune import autotune_batch
        a = time.perf_counter()
        cfg, tuned_ms, hist = autotune_batch(ctx, reqs, repeats=1, passes=1, log=log)
        tune = {"seconds": round(time.perf_counter() - a, 1), "candidates": len(hist), "best_ms": round(tuned_ms, 1)}
    b = Batch(ctx, reqs)
    if tune is not None:
        b.set_config(**cfg)

    def batch_rays():
        return sum(int(o["ray_counts"][0].item()) + int(o["ray_counts"][1].item()) for o in b.outputs)

    for _ in range(max(1, args.warmup)):  # the first render captures the graph
        b.render()
    torch.cuda.synchronize()
    times, rays = [], []
    for _ in range(args.steps):
        torch.cuda.synchronize()
        a = time.perf_counter()
        b.render()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - a)
        rays.append(batch_rays())
        log(f"batch: {times[-1] * 1e3:.1f} ms, {rays[-1] / 1e6:.1f} M rays")
    value = sum(rays) / sum(times) / 1e6
    ms = sum(times) / len(times) * 1e3

    # the same requests one after another (st_render, four wave slots each)
    outs = [None] * len(reqs)
    for i, (s, v, p) in enumerate(reqs):
        outs[i] = TestRenderer(s).render(v, p)
    torch.cuda.synchronize()
    seq = []
    for _ in range(max(1, min(args.steps, 2))):
        torch.cuda.synchronize()
        a = time.perf_counter()
        for i, (s, v, p) in enumerate(reqs):
            TestRenderer(s).render(v, p, out=outs[i])
        torch.cuda.synchronize()
        seq.append(time.perf_counter() - a)
    seq_ms = min(seq) * 1e3
    same = all(torch.equal(outs[i]["hdr"], b.outputs[i]["hdr"]) for i in range(len(reqs)))
    log(f"sequential st_render: {seq_ms:.1f} ms; batch images equal: {same}")

    # exclusive per-kernel times and algorithmic bytes, summed over requests
    totals = {}
    serial_ms = 0.0
    for (s, v, p), (kind, side, spp) in zip(reqs, mix):
        cloud = kind == "cloud"
        r = TestRenderer(s)
        cp = abi.default_params(side, side, max_depth=p.max_depth, spp=1, flags=p.flags | abi.ST_FLAG_COUNT_TRAVERSAL,
                                rr_start_depth=p.rr_start_depth)
        first = r.render(v, cp)
        torch.cuda.synchronize()
        tc = ctx.timings()
        prim_hits = int((first["visibility"][..., 0] != -1).sum().item())
        per_ray = (tc["closest_nodes"] / max(1, tc["closest_rays"]), tc["closest_tris"] / max(1, tc["closest_rays"]),
                   tc["shadow_nodes"] / max(1, tc["shadow_rays"]), tc["shadow_tris"] / max(1, tc["shadow_rays"]))
        sp = abi.default_params(side, side, max_depth=p.max_depth, spp=spp, flags=p.flags | abi.ST_FLAG_SERIAL_WAVES,
                                rr_start_depth=p.rr_start_depth)
        a = time.perf_counter()
        r.render(v, sp)
        torch.cuda.synchronize()
        serial_ms += (time.perf_counter() - a) * 1e3
        kx, ts = ctx.kernel_times(), ctx.timings()
        algo, names = algo_bytes(ts, per_ray, cloud, prim_hits, side * side * spp, p, not cloud)
        for k, nbytes in algo.items():
            ms_k, n_k = kx[k]
            if n_k == 0 or ms_k <= 0:
                continue
            t = totals.setdefault(names.get(k, k), [0.0, 0, 0.0])
            t[0] += ms_k
            t[1] += n_k
            t[2] += nbytes
    kernels = {k: {"ms_per_batch": round(t[0], 3), "launches": t[1], "avg_launch_ms": round(t[0] / t[1], 4),
                   "bytes_per_launch": round(t[2] / t[1]), "frac": round(t[2] / (t[0] / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
               for k, t in totals.items()}
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_batch"])
    d = kernels[dom]
    assert sum(k["ms_per_batch"] for k in kernels.values()) <= serial_ms, (kernels, serial_ms)
    achieved = d["bytes_per_launch"] / (d["avg_launch_ms"] / 1e3) / 1e9
    roofline = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "bytes_per_launch": d["bytes_per_launch"], "avg_launch_ms": d["avg_launch_ms"],
                "timing": "HIP events around each launch, every request rendered once more with ST_FLAG_SERIAL_WAVES "
                          f"({serial_ms:.1f} ms for all requests), summed over requests",
                "kernels": kernels}
    cpu = None
    if not args.no_cpu_baseline:
        # the oracle on a 512^2 spheres request of the mix, bounded like the headline's
        s_side = 512
        ns = argparse.Namespace(width=s_side, height=s_side, cpu_seconds=args.cpu_seconds)
        cpu = cpu_baseline(sd["spheres"], scenes.spheres_view(s_side, s_side), ns, 8, head, 3, log)
    px_spp = sum(side * side * spp for _, side, spp in mix)
    line = {
        "metric": "Mrays/s batch of 32 uneven requests", "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded scene generators)",
        "config": {"workload": "32 requests (seed 7): square 128-1024 px a side, 1/4/16/64 spp; spheres requests as the "
                               "headline (1,004,098 tris, maxPathLength 8, RR from 3, alias-sampled sky, ReSTIR DI with "
                               "reuse), every fourth on the 256^3 cloud (maxPathLength 16, delta + ratio tracking); one "
                               "st_batch_render (one hipGraph launch over 4 streams)",
                   "requests": [f"{k} {s}^2 {n}spp" for k, s, n in mix], "pixel_samples": px_spp,
                   "batch_config": b.config(), "autotune": tune, "sequential_st_render_ms": round(seq_ms, 2),
                   "batch_speedup_vs_sequential": round(seq_ms / ms, 3), "images_equal_sequential": bool(same),
                   "samples_per_px_per_s": round(px_spp / (ms / 1e3), 1)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    b.close()
    print(json.dumps(line), flush=True)


def bench_adjoint(args, log):
    """One step = one gradient: st_adjoint over 16 cameras (512x512, 8 spp,
    maxPathLength 8, RR from 3, Lambertian, fp32 NEE) of a random linear
    image loss, with respect to the per-triangle albedo and the 64^3 density
    grid.  The 16 forward renders (st_render, same arguments) are timed
    beside it; roofline figures come from one more adjoint per camera with
    ST_FLAG_SERIAL_WAVES (exclusive HIP-event times, both passes)."""
    import numpy as np
    import torch
    from stratum2_amd import abi, scenes
    from stratum2_amd.renderer import Context, Scene, TestRenderer
    W = H = 512
    spp, depth, ncam = 8, 8, 16
    ctx = Context(0)
    sd, offs = scenes.cornell_cloud(seed=0, n=64)
    sc = Scene(ctx, sd)
    T = int(sc.stats()["triangles"])
    rng = np.random.default_rng(0)
    alb = torch.from_numpy(rng.uniform(0.2, 0.9, (T, 3)).astype(np.float32)).cuda()
    sc.set_triangle_albedo(alb)
    views = scenes.cornell_cameras(W, H, ncam)
    flags = abi.ST_FLAG_DEFAULT | abi.ST_FLAG_LAMBERTIAN | abi.ST_FLAG_FLOAT_NEE
    params = [abi.default_params(W, H, max_depth=depth, spp=spp, flags=flags, rr_start_depth=3, seed=17 * i)
              for i in range(ncam)]
    dh = torch.from_numpy(rng.uniform(0.5, 1.5, (H, W, 4)).astype(np.float32)).cuda()
    words = sd.volumes[0].nbytes // 4
    ga = torch.zeros((T, 3), dtype=torch.float32, device="cuda")
    gg = torch.zeros(words, dtype=torch.float32, device="cuda")
    r = TestRenderer(sc)
    outs = [None] * ncam

    def forward():
        for i in range(ncam):
            outs[i] = r.render(views[i], params[i], out=outs[i])

    def gradient():
        for i in range(ncam):
            sc.adjoint(views[i], params[i], dh, ga, gg, 0)

    for _ in range(max(1, args.warmup)):
        forward()
        gradient()
    torch.cuda.synchronize()
    fwd, adj = [], []
    for _ in range(args.steps):
        torch.cuda.synchronize()
        a = time.perf_counter()
        forward()
        torch.cuda.synchronize()
        fwd.append(time.perf_counter() - a)
        ga.zero_()
        gg.zero_()
        torch.cuda.synchronize()
        a = time.perf_counter()
        gradient()
        torch.cuda.synchronize()
        adj.append(time.perf_counter() - a)
        log(f"forward {fwd[-1] * 1e3:.1f} ms, gradient {adj[-1] * 1e3:.1f} ms ({adj[-1] / fwd[-1]:.2f}x)")
    fwd_ms, adj_ms = sum(fwd) / len(fwd) * 1e3, sum(adj) / len(adj) * 1e3
    gv = gg[torch.as_tensor(offs.reshape(-1) // 4, device="cuda")]
    assert torch.isfinite(ga).all() and torch.isfinite(gg).all() and float(gv.abs().sum()) > 0

    # exclusive kernel times of the adjoint (both passes) and its ray counts;
    # per-ray grid reads from a 1-spp counting render of each camera
    totals, serial_ms = {}, 0.0
    for i in range(ncam):
        cp = abi.default_params(W, H, max_depth=depth, spp=1, flags=flags | abi.ST_FLAG_COUNT_TRAVERSAL, rr_start_depth=3,
                                seed=params[i].seed)
        r.render(views[i], cp)
        torch.cuda.synchronize()
        tc = ctx.timings()
        per_ray = (tc["closest_nodes"] / max(1, tc["closest_rays"]), 0.0, tc["shadow_nodes"] / max(1, tc["shadow_rays"]), 0.0)
        sp = abi.default_params(W, H, max_depth=depth, spp=spp, flags=flags | abi.ST_FLAG_SERIAL_WAVES, rr_start_depth=3,
                                seed=params[i].seed)
        a = time.perf_counter()
        sc.adjoint(views[i], sp, dh, ga, gg, 0)
        torch.cuda.synchronize()
        serial_ms += (time.perf_counter() - a) * 1e3
        kx, ts = ctx.kernel_times(), ctx.timings()
        algo, names = algo_bytes(ts, per_ray, True, 0, W * H * spp, sp, False)
        for k, nbytes in algo.items():
            ms_k, n_k = kx[k]
            if n_k == 0 or ms_k <= 0:
                continue
            t = totals.setdefault(names.get(k, k), [0.0, 0, 0.0])
            t[0] += ms_k
            t[1] += n_k
            t[2] += nbytes
    kernels = {k: {"ms_per_gradient": round(t[0], 3), "launches": t[1], "avg_launch_ms": round(t[0] / t[1], 4),
                   "bytes_per_launch": round(t[2] / t[1]), "frac": round(t[2] / (t[0] / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
               for k, t in totals.items()}
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_gradient"])
    d = kernels[dom]
    assert sum(k["ms_per_gradient"] for k in kernels.values()) <= serial_ms, (kernels, serial_ms)
    achieved = d["bytes_per_launch"] / (d["avg_launch_ms"] / 1e3) / 1e9
    roofline = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "bytes_per_launch": d["bytes_per_launch"], "avg_launch_ms": d["avg_launch_ms"],
                "timing": "HIP events around each launch of one more adjoint per camera with ST_FLAG_SERIAL_WAVES "
                          f"({serial_ms:.1f} ms for 16 cameras), both passes; the gradient atomics are not counted",
                "kernels": kernels}
    cpu = None
    if not args.no_cpu_baseline:
        cpu = adjoint_cpu_baseline(sd, views[0], alb.cpu().numpy(), dh.cpu().numpy(), args.cpu_seconds, depth, flags, log)
    line = {
        "metric": "adjoint ms/grad", "value": round(adj_ms, 2), "unit": "ms", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(adj_ms, 2), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded Cornell box and fBm density)",
        "config": {"workload": f"gradient of a linear image loss over {ncam} cameras at {W}x{H}, {spp} spp, maxPathLength "
                               f"{depth} (RR from 3), Lambertian, fp32 NEE, with respect to {T} per-triangle albedos and "
                               "a 64^3 density grid (Cornell box + fBm cloud, sigma 6, albedo 0.8, g 0.2)",
                   "cameras": ncam, "width": W, "height": H, "spp": spp, "max_path_length": depth,
                   "forward_ms": round(fwd_ms, 2), "gradient_over_forward": round(adj_ms / fwd_ms, 3)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def cache_scene(name):
    """(scene, view, params(spp, seed)) of the time-to-RMSE comparison: the
    volume scene at 1024x1024 (maxPathLength 16, roulette everywhere) or the
    1M-triangle scene at 1920x1080 with the headline frame's flags."""
    from stratum2_amd import abi, scenes
    if name == "cloud":
        sd = scenes.cloud(seed=0)
        W = H = 1024
        flags = abi.ST_FLAG_DEFAULT | abi.ST_FLAG_ENV_ALIAS

        def mk(spp, seed, extra=0):
            return abi.default_params(W, H, max_depth=16, spp=spp, seed=seed, flags=flags | extra, rr_start_depth=0)
        return sd, scenes.cloud_view(W, H), mk
    sd = scenes.spheres_field(seed=1)
    W, H = 1920, 1080
    flags = abi.ST_FLAG_DEFAULT | abi.ST_FLAG_ENV_ALIAS | abi.ST_FLAG_RESTIR_DI | abi.ST_FLAG_RESTIR_DI_REUSE

    def mk(spp, seed, extra=0):
        return abi.default_params(W, H, max_depth=8, spp=spp, seed=seed, flags=flags | extra, rr_start_depth=3,
                                  candidate_samples=32, max_m=3.0, spatial_neighbours=5, spatial_radius=30.0)
    return sd, scenes.spheres_view(W, H), mk


# Adam's learning rate for the online-trained cache: 1e-2 left the frozen
# cache's image mean 34% high on the cloud after 32 frames, 1e-3 40% low (not
# yet converged), 3e-3 within 2% on both scenes; a moving average of the
# parameters lagged behind over so few frames
CACHE_LR = 3e-3


def cache_run(name, depth, frame_spp, total_spp, log):
    """Plain and cached progressive renders of one scene against a 1024-spp
    plain reference (seeds 10^6 on).  Plain: st_render calls of 16 spp;
    cached: st_render_cached frames of frame_spp spp, each training one Adam
    step on its 2^16 rows.  Only the render calls are timed (synchronised
    before and after); errors are computed outside the timed calls."""
    import torch
    from stratum2_amd.cache import RadianceCache
    from stratum2_amd.renderer import Context, Scene, TestRenderer
    sd, view, mk = cache_scene(name)
    ctx = Context(0)
    sc = Scene(ctx, sd)
    r = TestRenderer(sc)
    p0 = mk(1, 0)
    W, H = p0.width, p0.height
    ref = torch.zeros((H, W, 3), dtype=torch.float64, device="cuda")
    for i in range(16):
        ref += r.render(view, mk(64, 1_000_000 + 64 * i))["hdr"][..., :3].double()
    ref /= 16

    def err(acc, n):
        d = acc / n - ref
        return d.pow(2).mean().sqrt().item(), (d.pow(2) / (ref.pow(2) + 1e-2)).mean().item()

    P = 16
    r.render(view, mk(P, 0))
    acc = torch.zeros_like(ref)
    el, plain = 0.0, []
    for c in range(total_spp // P):
        torch.cuda.synchronize()
        a = time.perf_counter()
        o = r.render(view, mk(P, P * c))
        torch.cuda.synchronize()
        el += time.perf_counter() - a
        acc += o["hdr"][..., :3].double() * P
        plain.append((P * (c + 1), el) + err(acc, P * (c + 1)))
    lo, hi = sd.bounds()
    # one training path per `stride` paths, set from an untimed frame so that
    # about 2^16 of them reach the cache vertex (rows with weight) per frame
    stride = max(1, W * H * frame_spp // 65536)
    warm = RadianceCache(ctx, lo, hi, learning_rate=CACHE_LR)
    warm.render(sc, view, mk(frame_spp, 0), cache_depth=depth, train_stride=stride, train_steps=1)
    frac = warm.last_times()[5] / -(-W * H * frame_spp // stride)
    stride = max(1, int(W * H * frame_spp * frac // 65536))
    warm.render(sc, view, mk(frame_spp, 0), cache_depth=depth, train_stride=stride, train_steps=1)
    warm.close()
    cache = RadianceCache(ctx, lo, hi, learning_rate=CACHE_LR)
    acc = torch.zeros_like(ref)
    el, cached, times, losses = 0.0, [], [], []
    for f in range(total_spp // frame_spp):
        torch.cuda.synchronize()
        a = time.perf_counter()
        o = cache.render(sc, view, mk(frame_spp, frame_spp * f), cache_depth=depth, train_stride=stride, train_steps=1, frame=f)
        torch.cuda.synchronize()
        el += time.perf_counter() - a
        times.append(cache.last_times())
        losses.append(o["loss"])
        acc += o["hdr"][..., :3].double() * frame_spp
        cached.append((frame_spp * (f + 1), el) + err(acc, frame_spp * (f + 1)))
    # bias: the trained cache, frozen, over 256 more spp of other seeds
    acc = torch.zeros_like(ref)
    for f in range(256 // frame_spp):
        o = cache.render(sc, view, mk(frame_spp, 500_000 + frame_spp * f), cache_depth=depth, train_stride=stride, train_steps=0,
                         frame=10_000 + f)
        acc += o["hdr"][..., :3].double() * frame_spp
    frozen_rmse = err(acc, 256)[0]
    bias = ((acc / 256).mean() - ref.mean()).item() / ref.mean().item()
    # exclusive inference time: one more frame with its waves one after another
    from stratum2_amd import abi
    cache.render(sc, view, mk(frame_spp, 900_000, abi.ST_FLAG_SERIAL_WAVES), cache_depth=depth, train_stride=stride,
                 train_steps=0, frame=20_000)
    torch.cuda.synchronize()
    serial = cache.last_times()
    cache.close()
    target_spp = 128
    tp, ep = [(x[1], x[2]) for x in plain if x[0] == target_spp][0]
    hit = [x for x in cached if x[2] <= ep]
    t_c = hit[0][1] if hit else None
    inf_ms = sum(t[0] for t in times) / len(times)
    queries = sum(t[4] for t in times) / len(times)
    train_ms = sum(t[2] for t in times) / len(times)
    rows = sum(t[5] for t in times) / len(times)
    res = {"scene": name, "width": W, "height": H, "cache_depth": depth, "frame_spp": frame_spp, "train_stride": stride,
           "training_rows_per_frame": -(-W * H * frame_spp // stride), "training_rows_with_weight": round(rows),
           "plain_ms_per_spp": round(plain[-1][1] / plain[-1][0] * 1e3, 3),
           "cached_ms_per_spp": round(cached[-1][1] / cached[-1][0] * 1e3, 3),
           "target": f"plain RMSE at {target_spp} spp", "target_rmse": round(ep, 4), "plain_ms": round(tp * 1e3, 1),
           "cached_ms": round(t_c * 1e3, 1) if t_c else None, "cached_spp_at_target": hit[0][0] if hit else None,
           "time_ratio": round(t_c / tp, 3) if t_c else None,
           "rmse_plain": {x[0]: round(x[2], 4) for x in plain if x[0] in (16, 32, 64, 128, 256)},
           "rmse_cached": {x[0]: round(x[2], 4) for x in cached if x[0] in (16, 32, 64, 128, 256)},
           "relmse_plain_256": round(plain[-1][3], 4), "relmse_cached_256": round(cached[-1][3], 4),
           "frozen_cache_rmse_256spp": round(frozen_rmse, 4), "frozen_cache_relative_bias_of_mean": round(bias, 5),
           "inference_ms_per_frame": round(inf_ms, 3), "queries_per_frame": round(queries), "training_ms_per_frame": round(train_ms, 3),
           "serial_inference_ms": round(serial[0], 4), "serial_inference_launches": int(serial[1]),
           "serial_queries": round(serial[4]),
           "loss_first_last": [round(losses[0], 3), round(losses[-1], 3)]}
    log(f"cache {name}: {json.dumps(res)}")
    return res


def bench_cache(args, log):
    """Time-to-RMSE of the learned radiance cache: the volume scene (the
    value) and the 1M-triangle headline scene, paths ending in the cache at
    their second vertex (cache_depth 2), 8-spp frames with 2^16 training rows
    each; the time the cached form needs to reach the RMSE the plain tracer
    has after 128 spp, over the plain tracer's time to get there.  Roofline:
    the fused inference kernel, whose algorithmic bytes per query are the
    64-byte query, 16 levels x 8 corners x 4 bytes of table and the path's
    16-byte radiance read and written."""
    runs = [cache_run("cloud", 2, 8, 256, log), cache_run("spheres", 2, 8, 256, log)]
    c = runs[0]
    qbytes = 64 + 16 * 8 * 4 + 32
    # the serial frame's inference launches (one per wave) plus its one target launch per wave
    inf_s = c["serial_inference_ms"] / 1e3
    achieved = c["serial_queries"] * qbytes / inf_s / 1e9
    roofline = {"kernel": "k_rc_infer_paths", "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "bytes_per_query": qbytes, "queries_per_frame": c["serial_queries"],
                "avg_launch_ms": round(c["serial_inference_ms"] / max(1, c["serial_inference_launches"]), 4),
                "timing": "HIP events around the inference and training-target launches of one 8-spp frame rendered with "
                          "ST_FLAG_SERIAL_WAVES (no kernel overlap)",
                "mfma_tflops": round(c["serial_queries"] * 4 * 2 * 64 * 64 / inf_s / 1e12, 2)}
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cache_cpu_baseline(args.cpu_seconds, log)
    line = {
        "metric": "time-to-RMSE cached / plain", "value": c["time_ratio"], "unit": "ratio", "n_gpus": 1, "steps": 1,
        "warmup": 1, "ms_per_step": c["cached_ms"], "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 render, fp16 cache (fp32 accumulation)", "data": "synthetic (seeded scene generators)",
        "config": {"workload": "time for st_render_cached (paths end in the hash-grid + MLP radiance cache at their second "
                               "vertex; 8-spp frames, one Adam step (lr 3e-3) on about 2^16 training rows per frame) to "
                               "reach the RMSE "
                               "st_render has after 128 spp (16-spp calls), both against a 1024-spp st_render reference: "
                               "1024^2 256^3 fBm cloud (maxPathLength 16) as the value, 1920x1080 1M-triangle ReSTIR scene "
                               "beside it", "runs": runs},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def cache_cpu_baseline(seconds, log):
    """The oracle's or_render_cached (no network: the cached estimator's path
    work) on the volume scene at 1024x1024, 1 spp, cache depth 2, OpenMP over
    this host's share of cores; reported in ms per spp."""
    from oracle.oracle import OracleScene
    sd, view, mk = cache_scene("cloud")
    o = OracleScene(sd)
    threads = min(16, os.cpu_count() or 1)
    a = time.perf_counter()
    o.render_cached(view, mk(1, 0), 2, 128, 0, threads=threads)
    dt = time.perf_counter() - a
    log(f"cpu baseline: or_render_cached 1024x1024 x 1 spp in {dt:.2f}s on {threads} threads")
    return {"value": round(dt * 1e3, 1), "unit": "ms per spp", "cores": threads, "kind": "port",
            "sample": f"CPU oracle or_render_cached (cache answering 0), 1024x1024 cloud, 1 spp, cache depth 2 ({dt:.1f} s)"}


def adjoint_cpu_baseline(sd, view, alb, dh, seconds, depth, flags, log):
    """The oracle's or_adjoint (OpenMP over this host's share of cores) on
    one camera at 512x512, spp grown until it takes about a third of
    --cpu-seconds; reported in ms per gradient of 16 cameras at 8 spp."""
    from stratum2_amd import abi
    from oracle.oracle import OracleScene
    o = OracleScene(sd)
    o.set_triangle_albedo(alb)
    threads = min(16, os.cpu_count() or 1)
    words = sd.volumes[0].nbytes // 4
    spp = 1
    while True:
        p = abi.default_params(512, 512, max_depth=depth, spp=spp, flags=flags, rr_start_depth=3)
        a = time.perf_counter()
        o.adjoint(view, p, dh, albedo=True, grid_vol=0, grid_words=words, threads=threads)
        dt = time.perf_counter() - a
        if dt >= seconds / 3 or spp >= 8:
            break
        spp = min(8, spp * 2)
    per_grad = dt / spp * 8 * 16 * 1e3
    log(f"cpu baseline: or_adjoint 512x512 x {spp} spp in {dt:.2f}s on {threads} threads")
    return {"value": round(per_grad, 1), "unit": "ms", "cores": threads, "kind": "port",
            "sample": f"CPU oracle or_adjoint, one camera at 512x512 and {spp} spp ({dt:.1f} s), scaled to 16 cameras x 8 spp"}


def bench_svgf(args, log):
    """SVGF denoiser: one step = one st_denoise call on a resident 1920x1080
    frame (radiance, albedo, previous-frame uv, visibility, depth) rendered
    once by the library at 1 spp; the history carries over between steps as
    under a static camera."""
    import numpy as np
    import torch
    from stratum2_amd import abi, scenes
    from stratum2_amd.renderer import Context, Denoiser, Scene, TestRenderer
    W, H = args.width, args.height
    sd = scenes.spheres_field(seed=1)
    view = scenes.spheres_view(W, H)
    ctx = Context(0)
    sc = Scene(ctx, sd)
    fl = abi.ST_FLAG_DEFAULT | abi.ST_FLAG_ENV_ALIAS
    frame = TestRenderer(sc).render(view, abi.default_params(W, H, max_depth=8, spp=1, flags=fl, rr_start_depth=3),
                                    want_aux=True)
    torch.cuda.synchronize()
    settings = dict(atrous_iterations=5, filter_type=abi.ST_FILTER_ATROUS, history_limit=32, history_tap=1)
    den = Denoiser(ctx, W, H, **settings)
    out = torch.empty_like(frame["hdr"])
    den.denoise(view, frame, out=out)   # the first call starts the history
    for _ in range(args.warmup):
        den.denoise(view, frame, out=out)
    torch.cuda.synchronize()
    times, tims = [], []
    for _ in range(args.steps):
        torch.cuda.synchronize()
        a = time.perf_counter()
        den.denoise(view, frame, out=out)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - a)
        tims.append(den.timings())
        log(f"step: {times[-1] * 1e3:.3f} ms, passes {tims[-1]}")
    total_s = sum(times)
    px = W * H
    value = px * len(times) / total_s / 1e6
    ms = total_s / len(times) * 1e3
    # algorithmic bytes per pixel and launch: accumulation reads visibility 8,
    # depth 16, previous uv 8, radiance 16, albedo 16 and the history at the
    # reprojected pixel (visibility 8, depth 4, colour 16, moments 8) and writes
    # colour 16, moments 8, visibility 8, depth 4 and the geometry image 16; the
    # variance estimate reads colour 16, moments 8, visibility 8, depth 16,
    # geometry 16 and writes 16; a filter pass reads colour 16, geometry 16,
    # depth 16 and writes 16 (neighbours are counted once: they are L2 hits of
    # other pixels' loads).
    t = tims[-1]
    classes = {"k_svgf_accumulate": (t["accumulate_ms"], t["accumulate_launches"], 152 * px),
               "k_svgf_variance": (t["variance_ms"], t["variance_launches"], 80 * px),
               "k_svgf_atrous": (t["atrous_ms"], t["atrous_launches"], 64 * px * t["atrous_launches"])}
    dom = max(classes, key=lambda k: classes[k][0])
    dms, dl, dbytes = classes[dom]
    achieved = dbytes / (dms / 1e3) / 1e9 if dms > 0 else 0.0
    traffic, src = measured_traffic(args, dom)
    roofline = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": src,
                "bytes_per_launch": round(dbytes / max(1, dl)), "avg_launch_ms": round(dms / max(1, dl), 4),
                "passes_ms": {k: round(v[0], 4) for k, v in classes.items()}}
    cpu = None
    if not args.no_cpu_baseline:
        cpu = svgf_cpu_baseline(frame, W, H, den.params, log)
    line = {
        "metric": "Mpixels/s SVGF 1080p", "value": round(value, 2), "unit": "Mpixels/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (a rendered frame of the seeded scene)",
        "config": {"workload": f"{W}x{H} SVGF: temporal accumulation (reprojection, albedo demodulation, normal and "
                               "depth tests, history limit 32), variance estimate, 5 a-trous 5x5 passes",
                   "width": W, "height": H, **settings},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def svgf_cpu_baseline(frame, W, H, params, log, seconds=10.0):
    """The oracle's denoiser (single-threaded C) on the same frame: a history
    start, then calls with all passes, repeated for about `seconds`."""
    from stratum2_amd import scenes
    from oracle.oracle import OracleDenoiser
    host = {k: frame[k].cpu().numpy() for k in ("hdr", "albedo", "prev_uv", "visibility", "depth")}
    view = scenes.make_view(W, H)
    den = OracleDenoiser(W, H)
    args = (host["hdr"], host["albedo"], host["prev_uv"], host["visibility"], host["depth"], params)
    den.denoise(view, *args)
    n, a = 0, time.perf_counter()
    while True:
        den.denoise(view, *args)
        n += 1
        dt = time.perf_counter() - a
        if dt >= seconds or n >= 200:
            break
    log(f"cpu baseline: SVGF on {W}x{H}, {n} frames in {dt:.2f}s")
    return {"value": round(W * H * n / dt / 1e6, 4), "unit": "Mpixels/s", "cores": 1, "kind": "port",
            "sample": f"CPU oracle (C restatement of the svgf kernels, one thread), the same {W}x{H} frame, "
                      f"{n} calls after the history start in {dt:.1f} s"}


def measured_traffic(args, kernel):
    """HBM bytes per launch of `kernel` from the PMC summary committed for
    this workload (profiles/hbm_traffic_<workload>.json: FETCH_SIZE and
    WRITE_SIZE per dispatch, scaled by the copy calibration)."""
    if args.workload in ("cloud", "svgf"):
        name = args.workload
    else:
        name = ("spheres_light" if args.light_trace else "spheres_bdpt" if args.bdpt else
                "spheres_plain" if args.no_restir else "spheres_restir")
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{name}.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        rec = json.load(f)
    k = rec.get("kernels", {}).get(kernel)
    if not k:
        return None, None
    return int(k["hbm_bytes_per_dispatch"]), os.path.relpath(path, ROOT)


def copy_calibration(log):
    """A plain device-to-device copy of 256 MiB, three times, for scaling the
    PMC byte counters (each copy reads and writes exactly 268435456 bytes)."""
    import torch
    a = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log("copy calibration: 3 copies of 268435456 bytes done")


def cpu_baseline(sd, view, args, max_depth, flags, rr, log):
    """The oracle (C restatement of the same estimator: same flags, roulette
    start, candidates and reuse, OpenMP over this host's share of cores) on
    the same frame at a bounded number of samples, grown until it takes about
    a third of --cpu-seconds."""
    from stratum2_amd import abi
    from oracle.oracle import OracleScene
    o = OracleScene(sd)
    threads = min(16, os.cpu_count() or 1)  # this host's share of cores
    spp = 1
    while True:
        p = abi.default_params(args.width, args.height, max_depth=max_depth, spp=spp, flags=flags, rr_start_depth=rr)
        a = time.perf_counter()
        res = o.render(view, p, threads=threads, want_path_length=False)
        dt = time.perf_counter() - a
        nr = int(res["ray_counts"][0]) + int(res["ray_counts"][1])
        if dt >= args.cpu_seconds / 3 or spp >= 64:
            break
        spp = min(64, spp * max(2, int(args.cpu_seconds / 3 / max(dt, 1e-3)) + 1))
    log(f"cpu baseline: {args.width}x{args.height} x {spp} spp, {nr} rays in {dt:.2f}s on {threads} threads")
    return {"value": round(nr / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"CPU oracle (C restatement of this estimator, the same flags), the same "
                      f"{args.width}x{args.height} frame at {spp} spp, maxPathLength {max_depth}, {nr} rays in {dt:.1f} s"}


if __name__ == "__main__":
    main()
