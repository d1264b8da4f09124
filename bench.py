#!/usr/bin/env python3
"""Render throughput benchmark for libquasi_hip.so on one GPU.

Workload (default): the 1M-triangle soup (centroids U[-1,1]^3, sizes
log-uniform in [1e-3, 5e-2], albedo U[0,1], four quad lights of radiance 15),
binned-SAH BVH2 (32 bins, leaf <= 4), 1920x1080, 64 spp, depth 8, Russian
roulette from bounce 3, next-event estimation, FP32.

One step = one complete 64-spp frame from a cleared accumulation, rendered
through the plugin C ABI (Q_plugin_render) in launches of --spp-per-launch
samples.  Rays = BVH queries (camera + extension + shadow, misses included),
counted on the device.  The scene and BVH are built and uploaded before the
timed region (build time reported separately).

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mrays/s (+ % HBM roofline) at 1080p, batched vs plain; adjoint/forward time ratio"
HBM_PEAK_GBS = 8000.0
# VALU lane throughput of MI355X: the published FP32 vector peak, 157.3 TFLOP/s,
# counts an FMA lane-operation as 2 flops (256 CUs x 128 lanes x 2.4 GHz)
VALU_PEAK_TLANE = 157.3 / 2
COUNTERS = ROOT / "profiles" / "counters_soup1m_1080p.json"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads(hardware: int) -> tuple[int, str]:
    """Threads for the CPU baseline: every hardware thread, unless the
    machine's job limits (OMP_NUM_THREADS, the CPU affinity mask) give this
    process fewer; the reason is reported beside the figure."""
    limits = [hardware]
    why = "all hardware threads"
    try:
        aff = len(os.sched_getaffinity(0))
        if aff < min(limits):
            limits.append(aff)
            why = f"CPU affinity mask of this process ({aff} CPUs)"
    except (AttributeError, OSError):
        pass
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and 0 < int(omp) < min(limits):
        limits.append(int(omp))
        why = f"OMP_NUM_THREADS={omp}: this process's share of the host's cores"
    return min(limits), why


def cpu_baseline(mesh, args, seconds: float):
    """The scalar oracle (FP32, std::thread pool) on this host's cores over a
    bounded, evenly spread subset of the same workload's pixels."""
    from oracle import oracle as O
    hardware = O.hardware_concurrency()
    threads, why = cpu_threads(hardware)
    model = cpu_model()
    log(f"CPU baseline: {threads} threads of {hardware} ({why}); {model}")
    s = O.Scene(mesh)
    common = dict(max_depth=args.depth, rr_start=3, integrator=1, rng_mode=1, seed=0, precision=32)
    W, H = args.width, args.height

    def rays_of(c):
        return c.camera_rays + c.extension_rays + c.shadow_rays

    # one core: samples 0..3 of every 61st pixel
    c1 = s.render(W, H, 4, pixel_stride=61, threads=1, **common).counts
    one_core = rays_of(c1) / c1.seconds / 1e6
    # all cores: passes of 4 samples over every 3rd pixel until the budget is spent
    paths = rays = 0
    secs = 0.0
    offset = 0
    while offset < args.spp and (secs < seconds or offset == 0):
        c = s.render(W, H, 4, sample_offset=offset, pixel_stride=3, threads=threads, **common).counts
        paths += c.camera_rays
        rays += rays_of(c)
        secs += c.seconds
        offset += 4
    return dict(value=rays / secs / 1e6, unit="Mrays/s", cores=threads, kind="port", value_1core=one_core,
                host_threads=hardware, cpu_model=model, threads_reason=why,
                sample=(f"oracle FP32, {W}x{H} soup, samples 0..{offset - 1} of every 3rd pixel "
                        f"({paths} paths, {rays} rays, {secs:.1f} s on {threads} threads; 1-core figure from "
                        f"samples 0..3 of every 61st pixel; BVH build excluded)"))


def roofline_of(kernel: str, avg_ms: float, launches: int, alg_bytes: float) -> dict:
    """Roofline of the dominant kernel.  The roof is VALU lane throughput:
    lane-operations per launch (VALU wave instructions x 64 x the active-lane
    fraction, read by rocprofv3 on this workload and kept in profiles/) over
    the launch time measured here.  Beside it: the texture addresser's busy
    fraction and vector-memory instructions per launch (the node loads: the
    unit that bound the 128-byte-node traversal), the L2-to-fabric bytes per
    launch (FETCH_SIZE and WRITE_SIZE scaled by a copy kernel; HBM or
    Infinity Cache) against the HBM peak, the L2 hit rate, the wave
    occupancy, and the algorithmic bytes of the frozen traversal model (which
    count every node visit, so their rate is not bounded by HBM: a
    cache-amplification figure, not a roof)."""
    sec = avg_ms * 1e-3
    out = {"bound": "valu", "achieved": None, "peak": VALU_PEAK_TLANE, "unit": "T lane-op/s", "frac": None,
           "traffic": None, "kernel": kernel, "avg_launch_ms": round(avg_ms, 3), "launches": launches,
           "alg_bytes_per_launch": round(alg_bytes), "alg_bytes_rate_gbs": round(alg_bytes / sec / 1e9, 1)}
    try:
        c = json.loads(COUNTERS.read_text())["kernels"][kernel]
    except (OSError, ValueError, KeyError):
        log(f"no counters for {kernel} in {COUNTERS}: roofline without counter figures")
        return out
    lane_ops = c["valu_insts_per_dispatch"] * 64 * c["active_lane_frac"]
    achieved = lane_ops / sec / 1e12
    traffic = c["fetch_kib_per_dispatch"] * 1024 * c["fetch_scale"] + c["write_kib_per_dispatch"] * 1024 * c["write_scale"]
    out.update(achieved=round(achieved, 3), frac=round(achieved / VALU_PEAK_TLANE, 4), traffic=round(traffic),
               lane_ops_per_launch=round(lane_ops),
               valu_insts_per_launch=c["valu_insts_per_dispatch"], active_lane_frac=c["active_lane_frac"],
               valu_issue_frac=round(c["valu_insts_per_dispatch"] * 64 / sec / 1e12 / VALU_PEAK_TLANE, 4),
               fabric={"achieved_gbs": round(traffic / sec / 1e9, 1), "peak_gbs": HBM_PEAK_GBS,
                       "frac": round(traffic / sec / 1e9 / HBM_PEAK_GBS, 4),
                       "note": "FETCH_SIZE + WRITE_SIZE: L2-to-fabric bytes (HBM or Infinity Cache)"},
               l2_hit_rate=c.get("l2_hit_rate"), occupancy_waves_per_simd=c.get("occupancy_waves_per_simd"),
               ta_busy_frac=c.get("ta_busy_frac"), vmem_insts_per_launch=c.get("vmem_insts_per_dispatch"),
               counters_source=str(COUNTERS.relative_to(ROOT)))
    return out


def bench_batch(args):
    """64 scenes of uneven size (room + soup, 1k..500k triangles, 128^2..1024^2),
    32 spp, depth 6, NEE: every scene rendered one after another with the
    plain wavefront engine (non-persistent traversal, and the persistent
    default), then all of them as one batch (shared queues, captured pass
    graph, autotuned trace configuration).  Scene preparation
    (BVH builds, uploads) is outside both timed regions; every scene's batched
    beauty image is compared with its plain one."""
    import numpy as np
    from quasi_amd import abi, scenes
    from quasi_amd.plugin import HipRenderer, make_camera

    t0 = time.perf_counter()
    batch = scenes.batch_scenes(num_scenes=args.batch_scenes)
    ntri = sum(m.num_triangles for m, _, _, _ in batch)
    npix = sum(w * h for _, w, h, _ in batch)
    log(f"batch: {len(batch)} scenes, {ntri} triangles, {npix} pixels generated in {time.perf_counter() - t0:.1f} s")
    spp, depth = args.batch_spp, 6
    common = dict(max_depth=depth, rr_start=3, rng_mode=1, integrator=1, engine=abi.Q_ENGINE_WAVEFRONT,
                  bvh_width=args.bvh_width, block_size=args.block, queue_chunk=args.chunk)

    def rays_of(st):
        return st.rays_camera + st.rays_extension + st.rays_shadow

    plain_s = 0.0
    plain_rays = 0
    plain_imgs = []
    per_scene = []   # (triangles, pixels, seconds, rays) of the persistent engine, scene by scene
    prog_s = 0.0
    prog_rays = 0
    flat_s = 0.0
    flat_rays = 0
    # the plain wavefront tracer: no persistent waves, one thread per queue slot
    flat = dict(common, bvh_width=4, block_size=128, queue_chunk=abi.Q_QUEUE_FLAT)
    with HipRenderer(64, 64, device=0) as r:
        for i, (mesh, w, h, seed) in enumerate(batch):
            r.set_scene(mesh)
            r.resize(w, h)
            c = mesh.camera
            r.camera = make_camera(c["position"], c["target"], c["up"], c["fov"])
            if i == 0:          # untimed warm-up of the kernels
                for kw in ((flat,) if args.batch_flat else ()) + (common,):
                    r.set_params(spp_per_render=spp, seed=seed, samples_per_pass=0, flags=0, **kw)
                    r.render(1)
                    r.synchronize()
            if args.batch_flat:
                r.set_params(spp_per_render=spp, seed=seed, samples_per_pass=0, flags=0, **flat)
                r.reset_stats()
                r.synchronize()
                t0 = time.perf_counter()
                r.render(1)
                r.synchronize()
                flat_s += time.perf_counter() - t0
                flat_rays += rays_of(r.stats())
            r.set_params(spp_per_render=spp, seed=seed, samples_per_pass=0, flags=0, **common)
            r.reset_stats()
            r.synchronize()
            t0 = time.perf_counter()
            r.render(1)
            r.synchronize()
            dt = time.perf_counter() - t0
            plain_s += dt
            nr = rays_of(r.stats())
            plain_rays += nr
            per_scene.append((mesh.num_triangles, w * h, dt, nr))
            plain_imgs.append(r.readback_aov()["beauty"])
            if args.batch_progressive:
                # the reference host's usage: one sample per Q_plugin_render call
                r.set_params(spp_per_render=1, seed=seed, samples_per_pass=0, flags=0, **common)
                r.reset_stats()
                r.synchronize()
                t0 = time.perf_counter()
                r.render(spp)
                r.synchronize()
                prog_s += time.perf_counter() - t0
                prog_rays += rays_of(r.stats())
        log(f"batch: plain sequential {plain_s:.3f} s, {plain_rays / plain_s / 1e6:.1f} Mrays/s")
        if flat_s:
            log(f"batch: plain non-persistent sequential {flat_s:.3f} s, {flat_rays / flat_s / 1e6:.1f} Mrays/s")

    with HipRenderer(64, 64, device=0) as r:
        r.set_params(spp_per_render=1, seed=0, samples_per_pass=args.batch_samples_per_pass, flags=0, **common)
        t0 = time.perf_counter()
        r.set_batch(batch)
        prep_s = time.perf_counter() - t0
        # probe with one pass of the size the timed render uses
        probe = args.batch_samples_per_pass or 4
        tune = r.autotune(abi.Q_AUTOTUNE_BATCH, probe_spp=probe, flags=abi.Q_AUTOTUNE_NO_STORE)
        best = tune.entries[tune.best]
        log(f"batch: prepared in {prep_s:.1f} s; autotuned width {best.bvh_width} block {best.block_size} "
            f"chunk {best.queue_chunk} ({best.ms:.2f} ms per probe)")
        # the tuner's probes warmed the kernels; it leaves the chosen pass graph
        # captured and the accumulation empty
        r.reset_stats()
        r.synchronize()
        t0 = time.perf_counter()
        r.render_batch(spp)
        r.synchronize()
        batched_s = time.perf_counter() - t0
        batched_rays = rays_of(r.stats())
        err = 0.0
        for i in range(len(batch)):
            img = r.readback_batch(i)["beauty"]
            ref = plain_imgs[i]
            err = max(err, float(np.abs(img - ref).sum() / max(np.abs(ref).sum(), 1e-30)))
    # the single-launch form of the batch: one persistent grid per captured
    # pass runs every path of every scene to its end (Q_ENGINE_PERSISTENT)
    single_s = 0.0
    single_rays = 0
    single_err = 0.0
    single_best = None
    if args.batch_single_launch:
        with HipRenderer(64, 64, device=0) as r:
            sl = dict(common, engine=abi.Q_ENGINE_PERSISTENT, block_size=256, queue_chunk=1)
            r.set_params(spp_per_render=1, seed=0, samples_per_pass=args.batch_samples_per_pass, flags=0, **sl)
            r.set_batch(batch)
            tune = r.autotune(abi.Q_AUTOTUNE_BATCH, probe_spp=probe, flags=abi.Q_AUTOTUNE_NO_STORE)
            single_best = tune.entries[tune.best]
            r.reset_stats()
            r.synchronize()
            t0 = time.perf_counter()
            r.render_batch(spp)
            r.synchronize()
            single_s = time.perf_counter() - t0
            single_rays = rays_of(r.stats())
            for i in range(len(batch)):
                img = r.readback_batch(i)["beauty"]
                ref = plain_imgs[i]
                single_err = max(single_err, float(np.abs(img - ref).sum() / max(np.abs(ref).sum(), 1e-30)))
        log(f"batch: single-launch {single_s:.3f} s, {single_rays / single_s / 1e6:.1f} Mrays/s "
            f"(width {single_best.bvh_width} block {single_best.block_size}, "
            f"{single_best.samples_per_pass} samples per pass), max rel-L1 vs plain {single_err:.2e}")
    plain = plain_rays / plain_s / 1e6
    batched = batched_rays / batched_s / 1e6
    # where the plain time goes, by scene size: triangle decades and resolutions
    by_class = {}
    for nt, npx, dt, nr in per_scene:
        key = f"{'1k-10k' if nt < 10_000 else '10k-100k' if nt < 100_000 else '100k-500k'} tris, {int(npx ** 0.5)}^2"
        c = by_class.setdefault(key, {"scenes": 0, "seconds": 0.0, "rays": 0})
        c["scenes"] += 1
        c["seconds"] += dt
        c["rays"] += nr
    for c in by_class.values():
        c["mrays"] = round(c["rays"] / c["seconds"] / 1e6, 1)
        c["seconds"] = round(c["seconds"], 4)
    log(f"batch: batched {batched_s:.3f} s, {batched:.1f} Mrays/s, speedup {batched / plain:.3f}, "
        f"max rel-L1 vs plain {err:.2e}")
    return {"workload": f"{len(batch)} scenes (room + soup, {ntri} triangles, {npix} pixels), {spp} spp, "
                        f"depth {depth}, RR from 3, NEE",
            "plain_mrays": round(plain, 3), "batched_mrays": round(batched, 3),
            "plain_by_class": dict(sorted(by_class.items())),
            "speedup": round(batched / plain, 4), "plain_s": round(plain_s, 4), "batched_s": round(batched_s, 4),
            "plain_rays": int(plain_rays), "batched_rays": int(batched_rays),
            "samples_per_pass": args.batch_samples_per_pass,
            "autotuned": {"bvh_width": best.bvh_width, "block_size": best.block_size,
                          "queue_chunk": best.queue_chunk, "samples_per_pass": best.samples_per_pass},
            "single_launch_mrays": round(single_rays / single_s / 1e6, 3) if single_s else None,
            "single_launch_speedup": round(single_rays / single_s / 1e6 / plain, 4) if single_s else None,
            "single_launch_max_rel_l1_vs_plain": single_err if single_s else None,
            "single_launch_autotuned": ({"bvh_width": single_best.bvh_width, "block_size": single_best.block_size,
                                         "samples_per_pass": single_best.samples_per_pass}
                                        if single_best is not None else None),
            "max_rel_l1_vs_plain": err,
            "plain_progressive_mrays": round(prog_rays / prog_s / 1e6, 3) if prog_s else None,
            "speedup_vs_progressive": round(batched / (prog_rays / prog_s / 1e6), 4) if prog_s else None,
            "plain_nonpersistent_mrays": round(flat_rays / flat_s / 1e6, 3) if flat_s else None,
            "speedup_vs_nonpersistent": round(batched / (flat_rays / flat_s / 1e6), 4) if flat_s else None}


def bench_adjoint(args):
    """Differentiable-rendering workload: icosphere of 9,680 triangles on a
    ground quad with two quad lights, 512 x 512, 32 spp, depth 4, NEE.  The
    forward pass renders the initial scene (32 spp from a cleared
    accumulation); the adjoint pass replays the same 32 samples and returns
    dL/d(albedo) and dL/d(vertex positions) of the L2 loss against a target
    render.  Medians of --adjoint-runs runs, wall clock (host calls included)
    and device time (hipEvents around the kernels)."""
    import numpy as np
    from quasi_amd import abi, scenes
    from quasi_amd.plugin import HipRenderer

    w = h = 512
    spp = 32
    init = scenes.diff_sphere(22, "init")
    target = scenes.diff_sphere(22, "target")
    common = dict(spp_per_render=spp, max_depth=4, rr_start=3, rng_mode=1, integrator=1, seed=1,
                  engine=abi.Q_ENGINE_WAVEFRONT, bvh_width=args.bvh_width, block_size=args.block,
                  queue_chunk=args.chunk, samples_per_pass=0)
    with HipRenderer(w, h, device=0) as r:
        for mesh in (target, init):
            r.set_scene(mesh)
            c = mesh.camera
            r.set_camera(c["position"], c["target"], c["up"], c["fov"])
            r.set_params(flags=abi.Q_FLAG_TIME_KERNELS, **common)
            r.render(1)
            if mesh is target:
                T = r.readback_aov()["beauty"]
        # the forward at its fastest traversal configuration for this scene
        tune = r.autotune(abi.Q_AUTOTUNE_SINGLE, probe_spp=spp, flags=abi.Q_AUTOTUNE_NO_STORE)
        best = tune.entries[tune.best]
        r.set_camera(init.camera["position"], init.camera["target"], init.camera["up"], init.camera["fov"])
        r.render(1)
        r.adjoint(T, spp, mode=abi.Q_ADJOINT_L2_TARGET)        # warm-up
        fwd_wall, fwd_dev, adj_wall, adj_dev = [], [], [], []
        for _ in range(args.adjoint_runs):
            r.reset_stats()
            r.synchronize()
            t0 = time.perf_counter()
            r.plugin.render(r.frame(camera_dirty=True))
            r.synchronize()
            fwd_wall.append(time.perf_counter() - t0)
            fwd_dev.append(r.stats().render_ms * 1e-3)
            t0 = time.perf_counter()
            ga, gv, res = r.adjoint(T, spp, mode=abi.Q_ADJOINT_L2_TARGET)
            adj_wall.append(time.perf_counter() - t0)
            adj_dev.append(res.kernel_ms * 1e-3)
        st = r.stats()
    med = lambda v: float(np.median(v))
    out = {"workload": f"icosphere {init.num_triangles} triangles ({init.vertices.shape[0]} vertices), "
                       f"{w}x{h}, {spp} spp, depth 4, NEE; gradients of an L2 image loss w.r.t. albedo and "
                       "vertex positions",
           "forward_ms": round(med(fwd_wall) * 1e3, 3), "adjoint_ms": round(med(adj_wall) * 1e3, 3),
           "forward_device_ms": round(med(fwd_dev) * 1e3, 3), "adjoint_device_ms": round(med(adj_dev) * 1e3, 3),
           "ratio": round(med(adj_wall) / med(fwd_wall), 4),
           "ratio_device": round(med(adj_dev) / med(fwd_dev), 4),
           "loss": res.loss, "paths": int(res.paths), "forward_rays": int(st.rays_camera + st.rays_extension
                                                                         + st.rays_shadow),
           "grad_albedo_l1": float(np.abs(ga).sum()), "grad_vertices_l1": float(np.abs(gv).sum()),
           "forward_traversal": {"bvh_width": best.bvh_width, "block_size": best.block_size,
                                 "queue_chunk": best.queue_chunk}}
    log(f"adjoint: forward {out['forward_ms']:.2f} ms, adjoint {out['adjoint_ms']:.2f} ms "
        f"(device {out['forward_device_ms']:.2f} / {out['adjoint_device_ms']:.2f} ms), ratio {out['ratio']:.3f}")
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--ntri", type=int, default=1_000_000)
    ap.add_argument("--spp-per-launch", type=int, default=64)
    ap.add_argument("--engine", choices=["megakernel", "wavefront"], default="wavefront")
    ap.add_argument("--samples-per-pass", type=int, default=0)
    ap.add_argument("--bvh-width", type=int, choices=[2, 4], default=4)
    ap.add_argument("--block", type=int, choices=[64, 128, 256], default=128)
    ap.add_argument("--chunk", type=int, choices=[1, 2, 3, 64, 256], default=3,
                    help="queue chunk: 1 per-lane fetch, 2 the same with the top BVH4 levels in LDS, 3 that over "
                         "quantised 64-byte BVH4 nodes, 64 or 256")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-scene batched-vs-plain measurement")
    ap.add_argument("--batch-scenes", type=int, default=64)
    ap.add_argument("--batch-spp", type=int, default=32)
    ap.add_argument("--batch-samples-per-pass", type=int, default=0, help="0: the plugin's automatic pass size")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward measurement")
    ap.add_argument("--adjoint-runs", type=int, default=5)
    ap.add_argument("--batch-flat", type=int, default=1,
                    help="also time the plain non-persistent wavefront tracer per scene (1/0)")
    ap.add_argument("--batch-progressive", type=int, default=1,
                    help="also time the plain engine at one sample per render call (the reference host's usage)")
    ap.add_argument("--batch-single-launch", type=int, default=1,
                    help="also time the batch in the single-launch form (one persistent grid per pass) (1/0)")
    ap.add_argument("--single-launch", type=int, default=1,
                    help="also time the headline frame with the single-launch engine (1/0)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}); "
              "multi-GPU sharding of this path is designed but not built", file=sys.stderr)
        sys.exit(2)
    if args.spp % args.spp_per_launch:
        sys.exit("--spp must be a multiple of --spp-per-launch")

    from quasi_amd import abi, scenes
    from quasi_amd.plugin import HipRenderer, make_camera

    t0 = time.perf_counter()
    mesh = scenes.soup(args.ntri, seed=1)
    log(f"scene: {mesh.num_triangles} triangles generated in {time.perf_counter() - t0:.1f} s")
    alg = json.loads((ROOT / "tests" / "golden" / "alg_bytes.json").read_text())["soup_1m_1080p"]

    W, H = args.width, args.height
    launches_per_step = args.spp // args.spp_per_launch
    with HipRenderer(W, H, device=0) as r:
        t0 = time.perf_counter()
        r.set_scene(mesh)
        build_s = time.perf_counter() - t0
        c = mesh.camera
        r.camera = make_camera(c["position"], c["target"], c["up"], c["fov"])
        engine = abi.Q_ENGINE_WAVEFRONT if args.engine == "wavefront" else abi.Q_ENGINE_MEGAKERNEL
        r.set_params(spp_per_render=args.spp_per_launch, max_depth=args.depth, rr_start=3, rng_mode=1,
                     integrator=1, seed=0, flags=abi.Q_FLAG_TIME_KERNELS, engine=engine,
                     samples_per_pass=args.samples_per_pass, bvh_width=args.bvh_width, block_size=args.block,
                     queue_chunk=args.chunk)
        st = r.stats()
        log(f"BVH: {st.bvh_nodes} nodes, depth {st.bvh_depth}, build+upload {st.bvh_build_ms:.0f} ms "
            f"(set_scene {build_s:.2f} s)")

        def step():
            for i in range(launches_per_step):
                f = r.frame(camera_dirty=(i == 0))
                r.plugin.render(f)

        for i in range(args.warmup):
            step()
            r.synchronize()
            log(f"warmup step {i + 1}/{args.warmup} done")
        r.reset_stats()
        r.synchronize()
        t0 = time.perf_counter()
        for i in range(args.steps):
            step()
        r.synchronize()
        elapsed = time.perf_counter() - t0
        st = r.stats()
        log(f"timed {args.steps} steps in {elapsed:.3f} s")

    single_launch = None
    if args.single_launch and args.engine == "wavefront":
        with HipRenderer(W, H, device=0) as r:
            r.set_scene(mesh)
            c = mesh.camera
            r.camera = make_camera(c["position"], c["target"], c["up"], c["fov"])
            r.set_params(spp_per_render=args.spp_per_launch, max_depth=args.depth, rr_start=3, rng_mode=1,
                         integrator=1, seed=0, flags=abi.Q_FLAG_TIME_KERNELS, engine=abi.Q_ENGINE_PERSISTENT,
                         samples_per_pass=args.samples_per_pass, bvh_width=args.bvh_width,
                         block_size=args.block if args.block in (128, 256) else 256, queue_chunk=1)

            def step_sl():
                for i in range(launches_per_step):
                    r.plugin.render(r.frame(camera_dirty=(i == 0)))

            for _ in range(args.warmup):
                step_sl()
                r.synchronize()
            r.reset_stats()
            r.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                step_sl()
            r.synchronize()
            sl_s = time.perf_counter() - t0
            ss = r.stats()
            sl_rays = ss.rays_camera + ss.rays_extension + ss.rays_shadow
            single_launch = {"mrays": round(sl_rays / sl_s / 1e6, 3), "ms_per_step": round(sl_s / args.steps * 1e3, 3),
                             "rays": int(sl_rays), "vs_wavefront": None}
        log(f"single-launch engine: {single_launch['mrays']:.1f} Mrays/s, {single_launch['ms_per_step']:.1f} ms/step")
    rays = {"camera": st.rays_camera, "extension": st.rays_extension, "shadow": st.rays_shadow}
    total_rays = sum(rays.values())
    value = total_rays / elapsed / 1e6
    kern = {st.kernels[i].name.decode(): st.kernels[i] for i in range(st.num_kernels)}
    dominant = "megakernel" if args.engine == "megakernel" else "wf_extend"
    k = kern[dominant]
    launches = int(k.launches)
    avg_ms = k.total_ms / max(1, launches)
    if dominant == "megakernel":
        bytes_rays = sum(rays[t] * alg[t]["bytes_per_ray"] for t in rays)
        bytes_fb = 32.0 * W * H * args.spp * args.steps
    else:   # the extend kernel traces the camera and extension rays only
        bytes_rays = sum(rays[t] * alg[t]["bytes_per_ray"] for t in ("camera", "extension"))
        bytes_fb = 0.0
    per_launch = (bytes_rays + bytes_fb) / max(1, launches)
    roofline = roofline_of(dominant, avg_ms, launches, per_launch)
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plain wavefront-config soup: {mesh.num_triangles} triangles, {W}x{H}, "
                               f"{args.spp} spp, depth {args.depth}, RR from 3, NEE",
                   "engine": args.engine, "spp_per_launch": args.spp_per_launch,
                   "bvh": f"binned-SAH BVH{args.bvh_width if args.engine == 'wavefront' else 2} over split "
                          f"triangle references",
                   "trace_block": args.block, "queue_chunk": args.chunk},
        "rays": rays,
        "rays_per_path": round(total_rays / (W * H * args.spp * args.steps), 4),
        "bvh_build_ms": round(st.bvh_build_ms, 1),
        "roofline": roofline,
        "kernel_ms": {n: round(v.total_ms, 3) for n, v in kern.items() if v.launches},
    }
    if single_launch:
        single_launch["vs_wavefront"] = round(single_launch["mrays"] / value, 4)
    out["single_launch"] = single_launch
    out["batch"] = None if args.no_batch else bench_batch(args)
    out["adjoint"] = None if args.no_adjoint else bench_adjoint(args)
    if not args.no_cpu:
        log("CPU baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(mesh, args, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
