"""Benchmark: closest-hit Mrays/s on the 3,407,872-triangle terrain scene.

One step = one frame of the closest-hit chain on one MI355X, inputs resident
in HBM: a fused 1920x1080 pinhole primary trace (coherent class; its 8x8
tiles start slowest first by the previous launch's per-tile step counts, as
a renderer tracing consecutive frames would run it; the plain order and a
moving camera are timed beside it under order_comparison), secondary-ray
generation from the primary hits (one cosine-weighted diffuse bounce and one
shadow ray per hit, stable wave64 compaction), and the closest-hit trace of
the diffuse rays (incoherent class), read off a device-side ray count. The
step's value is (primary + diffuse rays) / step time. The shadow any-hit rays
of the same frame are traced after each step and reported as their own class.

Roofline for the dominant kernel (the incoherent closest-hit trace):
algorithmic bytes per launch = sum over rays of 64*nodes + 36*tris + 32 + 16
(BVH4 node, SoA triangle, explicit ray in, hit out), with node and triangle
counts from the counted kernel, whose counts equal the CPU oracle's. The CPU
baseline is the oracle's multithreaded traversal on a bounded ray sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
METRIC = "Mrays/s closest-hit (coherent/incoherent, 2M tris); equal-time denoised PSNR gain"
HBM_PEAK_GBS = 8000.0
RAY_KEYS = ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def coherent_order_comparison(pkg, scenes, bvh, W, H, cam, out, stream, steps=20, orbit_frames=30):
    """Coherent-class launch times (ms, HIP events) in the plain tile order
    and in the cost order: on the bench camera, and on a camera orbiting the
    target by 1 and 3 degrees per frame, where each cost-ordered launch uses
    the previous frame's costs."""
    import math
    import torch

    def timed(cams, flags):
        evs = []
        for cc in cams:
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            bvh.trace_pinhole_closest(W, H, cc, *out, flags=flags)
            e1.record(stream)
            evs.append((e0, e1))
        torch.cuda.synchronize()
        return np.array([a.elapsed_time(b) for a, b in evs])

    res = {}
    timed([cam] * 3, 0)
    res["static_plain_ms"] = round(float(timed([cam] * steps, 0).mean()), 4)
    timed([cam] * 3, pkg.FLAG_COST_ORDER)
    res["static_cost_ordered_ms"] = round(float(timed([cam] * steps, pkg.FLAG_COST_ORDER).mean()), 4)
    eye, tgt = np.array(scenes.TERRAIN_EYE, np.float64), np.array(scenes.TERRAIN_TARGET, np.float64)
    for deg in (1.0, 3.0):
        cams = []
        for k in range(orbit_frames):
            a, d = math.radians(deg * k), eye - tgt
            e = tgt + np.array([d[0] * math.cos(a) - d[2] * math.sin(a), d[1], d[0] * math.sin(a) + d[2] * math.cos(a)])
            cc = scenes.look_at(tuple(e), scenes.TERRAIN_TARGET, W, H, 60.0)
            cams.append(pkg.camera(cc["o"], cc["f"], cc["r"], cc["u"], cc["scale_x"], cc["scale_y"]))
        timed(cams[:2], 0)
        plain = timed(cams, 0)[1:].mean()
        timed(cams[:1], pkg.FLAG_COST_ORDER)
        ordered = timed(cams, pkg.FLAG_COST_ORDER)[1:].mean()
        res[f"orbit_{deg:g}deg_per_frame"] = {"plain_ms": round(float(plain), 4),
                                              "cost_ordered_ms": round(float(ordered), 4)}
    # leave the bench camera's costs in place for the launches that follow
    timed([cam] * 2, pkg.FLAG_COST_ORDER)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-sample", type=int, default=1 << 22, help="max rays per class for the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (default launch configuration)")
    ap.add_argument("--no-pt", action="store_true", help="skip the path-tracer leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the many-scenes leg")
    ap.add_argument("--no-denoise", action="store_true", help="skip the denoiser leg")
    ap.add_argument("--batch-scenes", type=int, default=256)
    ap.add_argument("--ref-spp", type=int, default=4096, help="reference spp of the denoiser evaluation")
    ap.add_argument("--pmc-frame", action="store_true",
                    help="run two frames with the tuned launch configurations between two plain 512 MiB device "
                         "copies and exit (for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the ray workload would shard by image tiles, see DESIGN.md", file=sys.stderr)
        sys.exit(2)

    import torch
    pkg = importlib.import_module("vk-rt-cuda-raytracing_amd")
    scenes = importlib.import_module("vk-rt-cuda-raytracing_amd.scenes")
    dev = torch.device("cuda:0")
    W, H = args.width, args.height
    n = W * H

    t0 = time.time()
    tris = scenes.terrain_scene(seed=args.seed)
    log(f"scene: {tris.shape[0]} triangles ({time.time() - t0:.1f} s)")
    pkg.BVH.build(tris).destroy()          # warm-up: code objects and first allocations
    t0 = time.time()
    bvh = pkg.BVH.build(tris)
    build_s = time.time() - t0
    log(f"build: {build_s:.2f} s, {bvh.num_bvh4_nodes} BVH4 nodes, depth {bvh.depth}")

    stream = torch.cuda.Stream()
    bvh.set_stream(stream.cuda_stream)
    cam = scenes.terrain_camera(W, H)
    c = pkg.camera(cam["o"], cam["f"], cam["r"], cam["u"], cam["scale_x"], cam["scale_y"])
    light = scenes.TERRAIN_LIGHT

    with torch.cuda.stream(stream):
        pt = torch.empty(n, device=dev)
        pid = torch.empty(n, dtype=torch.int32, device=dev)
        pu = torch.empty(n, device=dev)
        pv = torch.empty(n, device=dev)
        D = {k: torch.empty(n, device=dev) for k in RAY_KEYS}
        S = {k: torch.empty(n, device=dev) for k in RAY_KEYS}
        src = torch.empty(n, dtype=torch.int32, device=dev)
        cnt = torch.zeros(1, dtype=torch.int32, device=dev)
        dt = torch.empty(n, device=dev)
        did = torch.empty(n, dtype=torch.int32, device=dev)
        du = torch.empty(n, device=dev)
        dv = torch.empty(n, device=dev)
        occ = torch.empty(n, dtype=torch.uint8, device=dev)

    def step(ev=None):
        if ev: ev[0].record(stream)
        # tiles ordered by the previous launch's per-tile step counts (slowest first)
        bvh.trace_pinhole_closest(W, H, c, pt, pid, pu, pv, flags=pkg.FLAG_COST_ORDER)
        if ev: ev[1].record(stream)
        bvh.gen_bounce_rays(W, H, c, pt, pid, 1234, light, D, S, src, cnt, flags=pkg.FLAG_TILE_ORDER)
        if ev: ev[2].record(stream)
        bvh.trace_closest_queued(n, cnt, D, dt, did, du, dv)
        if ev: ev[3].record(stream)
        bvh.trace_any_queued(n, cnt, S, occ)
        if ev: ev[4].record(stream)

    if args.pmc_frame:
        pmc_frame(pkg, bvh, step, dev, cnt)
        return

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    m = int(cnt.item())

    # Autotuner (untimed): LDS stack depth and resident workgroups per CU for
    # each ray class, measured on this frame's rays.
    tuning = {}
    if not args.no_tune:
        with torch.cuda.stream(stream):
            Dm = {k: v[:m] for k, v in D.items()}
            Sm = {k: v[:m] for k, v in S.items()}
            for name, cls, kw in (("coherent_closest", pkg.CLASS_COHERENT, dict(width=W, height=H, cam=c)),
                                  ("incoherent_closest", pkg.CLASS_INCOHERENT, dict(rays=Dm)),
                                  ("shadow_any", pkg.CLASS_ANY, dict(rays=Sm))):
                best, table = bvh.autotune(cls, reps=3, **kw)
                sched = SCHED_NAMES
                tuning[name] = {"schedule": sched[best[0]], "lds_depth": best[1], "blocks_per_cu": best[2],
                                "sweep_ms": {f"{sched[sc]}/{d}/{b}": round(ms, 4) for sc, d, b, ms in table}}
                log(f"autotune {name}: {sched[best[0]]} lds {best[1]} blocks/CU {best[2] or 'max'}  "
                    + " ".join(f"{sc}/{d}/{b}:{ms:.3f}" for sc, d, b, ms in table))
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
    hit_rate = float((pid >= 0).float().mean().item())

    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(5)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_wall = time.perf_counter()
    for k in range(args.steps):
        step(evs[k])
    torch.cuda.synchronize()
    wall_s = time.perf_counter() - t_wall
    seg = np.array([[e[i].elapsed_time(e[i + 1]) for i in range(4)] for e in evs])   # ms
    prim_ms, bounce_ms, diff_ms, shadow_ms = seg.mean(axis=0)
    step_ms = prim_ms + bounce_ms + diff_ms
    value = (n + m) / (step_ms * 1e3)
    diff_hit = float((did[:m] >= 0).float().mean().item())
    occ_rate = float(occ[:m].float().mean().item())
    log(f"primary {prim_ms:.3f} ms  bounce {bounce_ms:.3f} ms  diffuse {diff_ms:.3f} ms  shadow {shadow_ms:.3f} ms"
        f"  (wall {wall_s / args.steps * 1e3:.3f} ms/step incl. shadow)")

    # The coherent class without the cost order, and both orders on a moving
    # camera (orbiting the target, so the previous launch's costs are a
    # prediction, not the same frame's).
    with torch.cuda.stream(stream):
        order_cmp = coherent_order_comparison(pkg, scenes, bvh, W, H, c, (pt, pid, pu, pv), stream)
    log("coherent tile order: " + json.dumps(order_cmp))

    # Occupancy of each class's launch (VGPRs, LDS, workgroups and waves per
    # SIMD), read back after one more untimed launch of each.
    occupancy = {}
    with torch.cuda.stream(stream):
        for name, fn in (("coherent_closest", lambda: bvh.trace_pinhole_closest(W, H, c, pt, pid, pu, pv)),
                         ("incoherent_closest", lambda: bvh.trace_closest_queued(n, cnt, D, dt, did, du, dv)),
                         ("shadow_any", lambda: bvh.trace_any_queued(n, cnt, S, occ))):
            fn()
            occupancy[name] = bvh.launch_info()
    torch.cuda.synchronize()
    log("occupancy: " + "; ".join(f"{k}: {v['vgprs']} VGPRs, {v['lds_bytes']} B LDS, {v['waves_per_simd']} waves/SIMD"
                                  for k, v in occupancy.items()))

    # Incoherent class with the ray sort (octant + Morton keys, GPU radix
    # sort, gather, trace, scatter back) for comparison; reported, not in value.
    sorted_cls = {}
    with torch.cuda.stream(stream):
        Dm = {k: v[:m] for k, v in D.items()}
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        for _ in range(2):
            bvh.trace_closest(Dm, dt[:m], did[:m], du[:m], dv[:m], flags=pkg.FLAG_SORT_RAYS)
        e[0].record(stream)
        for _ in range(5):
            bvh.trace_closest(Dm, dt[:m], did[:m], du[:m], dv[:m], flags=pkg.FLAG_SORT_RAYS)
        e[1].record(stream)
        for _ in range(5):
            bvh.trace_closest(Dm, dt[:m], did[:m], du[:m], dv[:m])
        e[2].record(stream)
    torch.cuda.synchronize()
    srt_ms, uns_ms = e[0].elapsed_time(e[1]) / 5, e[1].elapsed_time(e[2]) / 5
    sorted_cls = {"rays": m, "sort_and_trace_ms": round(srt_ms, 4), "unsorted_trace_ms": round(uns_ms, 4),
                  "mrays_s_incl_sort": round(m / (srt_ms * 1e3), 1)}
    log(f"incoherent with ray sort: {srt_ms:.3f} ms (sort+trace) vs {uns_ms:.3f} ms unsorted")

    # The compressed 8-wide BVH on the same primary and diffuse rays (reported).
    cw_cls = {}
    with torch.cuda.stream(stream):
        t0 = time.time()
        ncw = bvh.build_cwbvh()
        cw_build_s = time.time() - t0
        Pp = {k: torch.empty(n, device=dev) for k in RAY_KEYS}
        pkg.gen_pinhole_rays(W, H, c, Pp, stream.cuda_stream)
        Dm = {k: v[:m] for k, v in D.items()}
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        for _ in range(2):
            bvh.trace_closest(Pp, pt, pid, pu, pv, flags=pkg.FLAG_CWBVH)
            bvh.trace_closest(Dm, dt[:m], did[:m], du[:m], dv[:m], flags=pkg.FLAG_CWBVH)
        e[0].record(stream)
        for _ in range(5):
            bvh.trace_closest(Pp, pt, pid, pu, pv, flags=pkg.FLAG_CWBVH)
        e[1].record(stream)
        for _ in range(5):
            bvh.trace_closest(Dm, dt[:m], did[:m], du[:m], dv[:m], flags=pkg.FLAG_CWBVH)
        e[2].record(stream)
    torch.cuda.synchronize()
    cwp, cwd = e[0].elapsed_time(e[1]) / 5, e[1].elapsed_time(e[2]) / 5
    cw_cls = {"nodes": ncw, "build_s": round(cw_build_s, 2),
              "coherent_closest_ms": round(cwp, 4), "coherent_mrays_s": round(n / (cwp * 1e3), 1),
              "incoherent_closest_ms": round(cwd, 4), "incoherent_mrays_s": round(m / (cwd * 1e3), 1)}
    log(f"CWBVH ({ncw} nodes, {cw_build_s:.2f} s): primary {cwp:.3f} ms, diffuse {cwd:.3f} ms")
    del Pp

    # Algorithmic bytes of the dominant kernel (incoherent closest hit) from
    # exact visit counts; the counted kernel's counts equal the oracle's.
    with torch.cuda.stream(stream):
        cn = torch.empty(n, dtype=torch.int32, device=dev)
        ct = torch.empty(n, dtype=torch.int32, device=dev)
        Dm = {k: v[:m] for k, v in D.items()}
        bvh.trace_closest_counted(Dm, dt[:m], did[:m], du[:m], dv[:m], cn[:m], ct[:m])
        pcn = torch.empty(n, dtype=torch.int32, device=dev)
        pct = torch.empty(n, dtype=torch.int32, device=dev)
        Pr = {k: torch.empty(n, device=dev) for k in RAY_KEYS}
        pkg.gen_pinhole_rays(W, H, c, Pr, stream.cuda_stream)
        bvh.trace_closest_counted(Pr, pt, pid, pu, pv, pcn, pct)
    torch.cuda.synchronize()
    d_nodes, d_tris = int(cn[:m].sum().item()), int(ct[:m].sum().item())
    p_nodes, p_tris = int(pcn.sum().item()), int(pct.sum().item())
    diff_bytes = 64 * d_nodes + 36 * d_tris + (32 + 16) * m
    achieved = diff_bytes / (diff_ms * 1e-3) / 1e9
    prim_bytes = 64 * p_nodes + 36 * p_tris + 16 * n

    # The reference boundary's host-pointer calls: tracePrimary (side x side
    # rays from a point looking +z, hit distances copied to a host array) and
    # traceDiffuse (host SoA rays in, host hit distances out). Their rates
    # include the PCIe copies and are reported beside the value, not in it.
    host_calls = {}
    with torch.cuda.stream(stream):
        side = 1440
        hit_host = np.zeros(side * side, np.float32)
        cam_o = (0.0, 2.0, -12.0)
        bvh.trace_primary(side, cam_o, hit_host)                 # warm-up
        reps = 5
        t0 = time.perf_counter()
        kern = [bvh.trace_primary(side, cam_o, hit_host) for _ in range(reps)]
        prim_wall = (time.perf_counter() - t0) / reps
        o_h = torch.stack([D[k][:m] for k in ("ox", "oy", "oz")], 1).cpu().numpy()
        d_h = torch.stack([D[k][:m] for k in ("dx", "dy", "dz")], 1).cpu().numpy()
        dh = np.zeros(m, np.float32)
        bvh.trace_diffuse(o_h, d_h, dh)
        t0 = time.perf_counter()
        kd = [bvh.trace_diffuse(o_h, d_h, dh) for _ in range(reps)]
        diff_wall = (time.perf_counter() - t0) / reps
    host_calls = {
        "tracePrimary": {"rays": side * side, "wall_ms": round(prim_wall * 1e3, 3),
                         "mrays_s_incl_pcie": round(side * side / prim_wall / 1e6, 1),
                         "mrays_s_reported": round(float(np.median(kern)), 1),
                         "hit_rate": round(float((hit_host < 1e29).mean()), 4)},
        "traceDiffuse": {"rays": m, "wall_ms": round(diff_wall * 1e3, 3),
                         "mrays_s_incl_pcie": round(m / diff_wall / 1e6, 1),
                         "mrays_s_reported": round(float(np.median(kd)), 1),
                         "note": "pageable host arrays, GPU octant+Morton sort included"},
    }
    log(f"host-pointer boundary: tracePrimary {host_calls['tracePrimary']['mrays_s_incl_pcie']} Mrays/s incl. PCIe "
        f"({host_calls['tracePrimary']['mrays_s_reported']} reported), traceDiffuse "
        f"{host_calls['traceDiffuse']['mrays_s_incl_pcie']} Mrays/s incl. PCIe "
        f"({host_calls['traceDiffuse']['mrays_s_reported']} reported)")

    cpu = None
    build = build_report(bvh, tris, build_s)
    if not args.no_cpu:
        cpu = cpu_baseline(bvh, Pr, D, m, args.cpu_sample, gpu_hits=((pid, pt), (did, dt)))
        build.update(cpu_build_baseline(bvh, tris))

    ic = tuning.get("incoherent_closest")
    prof_traffic = load_traffic(f"{ic['schedule']}/{ic['lds_depth']}/{ic['blocks_per_cu']}") if ic else None
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "terrain: 1024x1024-cell heightfield + 256 level-4 icospheres, 1920x1080 primary + "
                        "one diffuse bounce per hit (closest hit); shadow any-hit reported per class",
            "triangles": int(tris.shape[0]), "width": W, "height": H, "seed": args.seed,
            "camera": {"eye": scenes.TERRAIN_EYE, "target": scenes.TERRAIN_TARGET, "vfov_deg": 60},
            "light": light, "bvh": "binned-SAH BVH4, 16 bins, <=4 tris/leaf, FP16 bounds rounded outward",
            "isect": "watertight",
        },
        "classes": {
            "coherent_closest": {"rays": n, "ms": round(prim_ms, 4), "mrays_s": round(n / (prim_ms * 1e3), 1),
                                 "hit_rate": round(hit_rate, 4), "nodes_per_ray": round(p_nodes / n, 2),
                                 "tris_per_ray": round(p_tris / n, 2),
                                 "alg_GBps": round(prim_bytes / (prim_ms * 1e-3) / 1e9, 1),
                                 "tile_order": "cost: each XCD band's 8x8 tiles predicted at >= 48 steps first, "
                                               "slowest first, by the previous launch's per-tile step counts "
                                               "dilated by 2 tiles (CUDABVH_FLAG_COST_ORDER); the rest in strip order",
                                 "order_comparison": order_cmp},
            "incoherent_closest": {"rays": m, "ms": round(diff_ms, 4), "mrays_s": round(m / (diff_ms * 1e3), 1),
                                   "hit_rate": round(diff_hit, 4), "nodes_per_ray": round(d_nodes / max(m, 1), 2),
                                   "tris_per_ray": round(d_tris / max(m, 1), 2)},
            "bounce_generation": {"ms": round(bounce_ms, 4)},
            "incoherent_closest_sorted": sorted_cls,
            "cwbvh": cw_cls,
            "shadow_any": {"rays": m, "ms": round(shadow_ms, 4), "mrays_s": round(m / (shadow_ms * 1e3), 1),
                           "occluded": round(occ_rate, 4)},
        },
        "roofline": {
            "kernel": "k_trace closest-hit, explicit rays (incoherent diffuse class)",
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": prof_traffic,
            "launch_config": (f"{ic['schedule']}/{ic['lds_depth']}/{ic['blocks_per_cu']}" if ic else "default"),
            "alg_bytes_per_launch": diff_bytes, "kernel_ms": round(diff_ms, 4),
        },
        "cpu_baseline": cpu,
        "occupancy": occupancy,
        "host_pointer_calls": host_calls,
        "build_s": round(build_s, 3),
        "build": build,
        "tuning": tuning,
    }
    bvh.set_stream(None)
    bvh.destroy()
    del D, S, dt, did, du, dv, occ
    torch.cuda.empty_cache()
    if not args.no_pt:
        out["path_tracer"] = bench_path_tracer(pkg, scenes, dev, args)
    if not args.no_batch:
        out["batched_scenes"] = bench_batched(pkg, scenes, dev, args)
    if not args.no_denoise:
        out["denoiser"] = bench_denoiser(pkg, scenes, dev, args, out.get("path_tracer", {}).get("tuning", {}))
    print(json.dumps(out), flush=True)


def bench_path_tracer(pkg, scenes, dev, args):
    """Wavefront path tracer, instanced scene (64 instances of 4 meshes),
    1920x1080, 1 spp, depth 8, Russian roulette from ray 3; one sample's
    launch sequence replayed from a hipGraph. Rays = extension + shadow."""
    import torch
    sc = scenes.instanced_scene()
    t0 = time.time()
    blas = [pkg.BVH.build(m) for m in sc["meshes"]]
    gs = pkg.Scene(blas, sc["o2w"], sc["blas_of"], sc["albedo"])
    torch.cuda.synchronize()
    build_s = time.time() - t0
    W, H = 1920, 1080
    n = W * H
    cam = scenes.instanced_camera(W, H)
    c = pkg.camera(cam["o"], cam["f"], cam["r"], cam["u"], cam["scale_x"], cam["scale_y"])
    tuning = {}
    if not args.no_tune:
        R = {k: torch.empty(n, device=dev) for k in RAY_KEYS}
        pkg.gen_pinhole_rays(W, H, c, R)
        t = torch.empty(n, device=dev)
        inst = torch.empty(n, dtype=torch.int32, device=dev)
        pid = torch.empty(n, dtype=torch.int32, device=dev)
        gs.trace_closest(R, t, inst, pid)
        torch.cuda.synchronize()
        hit = pid >= 0
        m = int(hit.sum())
        o = torch.stack([R[k] for k in ("ox", "oy", "oz")])[:, hit]
        d = torch.stack([R[k] for k in ("dx", "dy", "dz")])[:, hit]
        P = o + t[hit] * d
        sun = torch.tensor(scenes.SUN_DIR, device=dev)
        sun = sun / sun.norm()
        S = {"ox": P[0] + 1e-3 * sun[0], "oy": P[1] + 1e-3 * sun[1], "oz": P[2] + 1e-3 * sun[2],
             "dx": torch.full((m,), float(sun[0]), device=dev), "dy": torch.full((m,), float(sun[1]), device=dev),
             "dz": torch.full((m,), float(sun[2]), device=dev), "tmin": torch.zeros(m, device=dev),
             "tmax": torch.full((m,), 1e30, device=dev)}
        S = {k: v.contiguous() for k, v in S.items()}
        # scattered rays for the later bounces: from the same hit points, random
        # directions in the hemisphere facing back towards the camera
        g = torch.Generator(device=dev)
        g.manual_seed(5)
        dirs = torch.randn(3, m, device=dev, generator=g)
        dirs = dirs / dirs.norm(dim=0, keepdim=True)
        dirs = torch.where((dirs * d).sum(dim=0, keepdim=True) > 0, -dirs, dirs)
        B = {"ox": P[0] + 1e-3 * dirs[0], "oy": P[1] + 1e-3 * dirs[1], "oz": P[2] + 1e-3 * dirs[2],
             "dx": dirs[0], "dy": dirs[1], "dz": dirs[2], "tmin": torch.zeros(m, device=dev),
             "tmax": torch.full((m,), 1e30, device=dev)}
        B = {k: v.contiguous() for k, v in B.items()}
        for name, cls, rays, cnt in (("closest", 0, R, n), ("shadow_any", 1, S, m), ("bounce_closest", 2, B, m)):
            best, table = gs.autotune(cls, cnt, rays, reps=3)
            sched = ("while-while", "batch", "batch-nearest-first", "refill")
            tuning[name] = {"schedule": sched[best[0]], "lds_depth": best[1], "blocks_per_cu": best[2],
                            "sweep_ms": {f"{sched[sc]}/{d}/{b}": round(ms, 4) for sc, d, b, ms in table}}
            log(f"pt autotune {name}: {sched[best[0]]} lds {best[1]} blocks/CU {best[2] or 'max'} "
                f"({min(x[3] for x in table):.3f} ms)")
        del R, S, B, t, inst, pid
    rgb = torch.empty(3 * n, device=dev)
    alb = torch.empty(3 * n, device=dev)
    nrm = torch.empty(3 * n, device=dev)
    dep = torch.empty(n, device=dev)
    if not args.no_tune:
        # queue size: below how many live paths one thread per path finishes the bounces
        best_tail, table = gs.autotune_path_tracer(c, pkg.path_params(W, H, spp=1, seed=1), rgb, alb, nrm, dep, reps=4)
        tuning["tail_paths"] = {"chosen": best_tail, "sweep_ms": {str(t): round(ms, 4) for t, ms in table}}
        log(f"pt autotune tail paths: {best_tail}  " + " ".join(f"{t}:{ms:.3f}" for t, ms in table))
    K = 16
    timed = {}
    for name, extra in (("unsorted", 0), ("octant_sorted", pkg.FLAG_SORT_RAYS)):
        # one graph capture per variant (sorting adds kernels), then K replayed samples
        p = pkg.path_params(W, H, spp=1, seed=1, tail_paths=-1)
        cnt = gs.path_trace(c, p, rgb, alb, nrm, dep, counts=True, flags=pkg.FLAG_GRAPH | extra)
        p = pkg.path_params(W, H, spp=K, seed=2, tail_paths=-1)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        gs.path_trace(c, p, rgb, alb, nrm, dep, flags=pkg.FLAG_GRAPH | extra)
        e1.record()
        torch.cuda.synchronize()
        timed[name] = e0.elapsed_time(e1) / K
    best = min(timed, key=timed.get)
    ms = timed[best]
    rays = int(cnt.sum())
    d = p.maxDepth
    log(f"path tracer: {ms:.3f} ms/sample ({best}; " + ", ".join(f"{k} {v:.3f}" for k, v in timed.items())
        + f"), {rays} rays/sample, {rays / (ms * 1e3):.1f} Mrays/s")
    gs.destroy()
    return {"workload": "64 instances of 4 heightfield meshes (10,082/49,928/199,712/500,000 tris), 1920x1080, "
                        "1 spp, depth 8, RR from ray 3, sun NEE + sky, hipGraph replay",
            "instanced_triangles": int(sum(len(sc["meshes"][b]) for b in sc["blas_of"])),
            "unique_triangles": int(sum(len(m) for m in sc["meshes"])),
            "ms_per_sample": round(ms, 3), "rays_per_sample": rays, "mrays_s": round(rays / (ms * 1e3), 1),
            "extension_ray_order": best,
            "ms_per_sample_by_order": {k: round(v, 3) for k, v in timed.items()},
            "extension_rays_per_bounce": [int(x) for x in cnt[:d]],
            "shadow_rays_per_bounce": [int(x) for x in cnt[d:]],
            "build_s": round(build_s, 2), "tuning": tuning}


def bench_batched(pkg, scenes, dev, args):
    """Many independent scenes: random-triangle scenes with log-uniform
    triangle counts in [1K, 200K] and ray counts in [64K, 1M], all BVHs built
    in one forest pass, all rays traced by one persistent launch (device work
    counter) replayed from a hipGraph. Aggregate Mrays/s, closest hit."""
    import torch
    specs = scenes.batch_specs(args.batch_scenes)
    t0 = time.time()
    tris = [scenes.random_triangles(nt, seed=5000 + k) for k, (nt, _) in enumerate(specs)]
    gen_s = time.time() - t0
    t0 = time.time()
    b = pkg.Batch(tris)
    torch.cuda.synchronize()
    build_s = time.time() - t0
    ntris = int(b.tri_offsets[-1])
    del tris
    R, roff = scenes.batch_workload_device(specs, dev)
    total = int(roff[-1])
    d_off = torch.from_numpy(roff).to(dev)
    t = torch.empty(total, device=dev)
    pid = torch.empty(total, dtype=torch.int32, device=dev)
    u = torch.empty(total, device=dev)
    v = torch.empty(total, device=dev)
    for _ in range(2):
        b.trace_closest(total, d_off, R, t, pid, u, v, flags=pkg.FLAG_GRAPH)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    K = 5
    e0.record()
    for _ in range(K):
        b.trace_closest(total, d_off, R, t, pid, u, v, flags=pkg.FLAG_GRAPH)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / K
    hit = float((pid >= 0).float().mean())
    log(f"batched: {len(specs)} scenes, {ntris} tris built in {build_s:.2f} s; {total} rays in {ms:.2f} ms "
        f"= {total / (ms * 1e3):.1f} Mrays/s (hit {hit:.3f})")
    b.destroy()
    return {"workload": f"{len(specs)} random-triangle scenes, tris log-uniform [1K,200K], rays log-uniform "
                        "[64K,1M] (half primary, half random), one forest build, one graph-replayed launch",
            "scenes": len(specs), "triangles": ntris, "rays": total, "build_s": round(build_s, 2),
            "scene_generation_s": round(gen_s, 2), "trace_ms": round(ms, 3),
            "mrays_s": round(total / (ms * 1e3), 1), "hit_rate": round(hit, 4)}


def denoiser_gflop(W, H):
    """Multiply-adds of one frame counted from the layer list (2 FLOP each)."""
    dn_layers = [(16, 32, 9, 0), (32, 32, 9, 0), (32, 64, 9, 1), (64, 64, 9, 1), (64, 128, 9, 2), (128, 128, 9, 2),
                 (128, 256, 9, 3), (256, 256, 9, 3), (384, 128, 9, 2), (128, 128, 9, 2), (192, 64, 9, 1),
                 (64, 64, 9, 1), (96, 32, 9, 0), (32, 32, 9, 0), (32, 3, 1, 0)]
    macs = sum(cin * cout * taps * (W >> lv) * (H >> lv) for cin, cout, taps, lv in dn_layers)
    return 2.0 * macs / 1e9


def bench_denoiser(pkg, scenes, dev, args, pt_tuning=None):
    """The learned denoiser at 1920x1080: inference time and TFLOP/s (BF16
    MFMA), and the equal-time quality gain over plain sample accumulation
    against a high-spp reference of the held-out bench view (PSNR of x/(1+x))."""
    import torch
    dt = importlib.import_module("vk-rt-cuda-raytracing_amd.denoise_train")
    W, H = 1920, 1080
    n = W * H
    dn = pkg.Denoiser(1920 * 1088, seed=0)
    wpath = dt.WEIGHTS
    trained = os.path.exists(wpath)
    if trained:
        dn.set_params(dt.load_weights(wpath))
    sc = scenes.instanced_scene(seed=2)
    r = dt.Renderer(sc, dev)
    sched = {"while-while": 0, "batch": 1, "batch-nearest-first": 2, "refill": 3}
    for cls, key in ((0, "closest"), (1, "shadow_any"), (2, "bounce_closest")):   # the path-tracer leg's tuned launches
        tcfg = (pt_tuning or {}).get(key)
        if tcfg:
            r.scene.set_trace_config(cls, sched[tcfg["schedule"]], tcfg["lds_depth"], tcfg["blocks_per_cu"])
    cam = scenes.instanced_camera(W, H)
    noisy, _ = r.render(cam, W, H, 1, 5, flags=pkg.FLAG_GRAPH)
    out = torch.empty(3 * n, device=dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    dn.denoise(W, H, noisy["rgb"], noisy["alb"], noisy["nrm"], noisy["dep"], out)
    torch.cuda.synchronize()
    K = 5
    e0.record()
    for _ in range(K):
        dn.denoise(W, H, noisy["rgb"], noisy["alb"], noisy["nrm"], noisy["dep"], out)
    e1.record()
    torch.cuda.synchronize()
    t_dn = e0.elapsed_time(e1) / K
    e0.record()
    for k in range(K):
        noisy, _ = r.render(cam, W, H, 1, 5, flags=pkg.FLAG_GRAPH)
    e1.record()
    torch.cuda.synchronize()
    t_pt = e0.elapsed_time(e1) / K
    gflop = denoiser_gflop(W, H)
    res = {"workload": "U-Net 32/64/128/256, 1920x1080, 16-channel BF16 input (10 used), FP32 accumulation",
           "inference_ms": round(t_dn, 3), "gflop_per_frame": round(gflop, 1),
           "achieved_tflops": round(gflop / t_dn, 1), "peak_tflops_bf16_dense": 2500.0,
           "mfma_frac": round(gflop / t_dn / 2500.0, 4), "path_trace_ms_per_spp": round(t_pt, 3),
           "trained_weights": os.path.relpath(wpath, ROOT) if trained else None}
    log(f"denoiser: {t_dn:.2f} ms/frame ({gflop / t_dn:.1f} TFLOP/s); path tracer {t_pt:.2f} ms/spp")
    if trained:
        ref, _ = r.render(cam, W, H, args.ref_spp, 999, flags=pkg.FLAG_GRAPH)
        k = max(1, int(np.ceil((t_pt + t_dn) / t_pt)))
        acc, _ = r.render(cam, W, H, k, 77, flags=pkg.FLAG_GRAPH)
        dn.denoise(W, H, noisy["rgb"], noisy["alb"], noisy["nrm"], noisy["dep"], out)
        torch.cuda.synchronize()
        refn = ref["rgb"].cpu().numpy().reshape(3, H, W)
        one = noisy["rgb"].cpu().numpy().reshape(3, H, W)
        den = out.cpu().numpy().reshape(3, H, W)
        accn = acc["rgb"].cpu().numpy().reshape(3, H, W)
        p1, pk, pd = dt.psnr(one, refn), dt.psnr(accn, refn), dt.psnr(den, refn)
        res.update({"reference_spp": args.ref_spp, "equal_time_spp": k, "psnr_1spp": round(p1, 2),
                    "psnr_equal_time_accumulation": round(pk, 2), "psnr_1spp_denoised": round(pd, 2),
                    "psnr_gain_db": round(pd - pk, 2),
                    "relmse_equal_time_accumulation": round(dt.rel_mse(accn, refn), 5),
                    "relmse_1spp_denoised": round(dt.rel_mse(den, refn), 5)})
        log(f"denoiser quality: 1spp {p1:.2f} dB, {k} spp {pk:.2f} dB, 1spp+denoiser {pd:.2f} dB "
            f"(gain {pd - pk:+.2f} dB vs equal time)")
    r.destroy()
    dn.destroy()
    return res


# Launch configurations the autotuner picks on the 1080p terrain frame on
# MI355X (schedule, LDS stack depth, workgroups per CU; 0 = as many as fit);
# for the incoherent class it alternates between close candidates from run to
# run, so each of them gets its own PMC frames.
SCHED_NAMES = ("batch", "while-while", "packet", "refill")
PMC_FRAME_CONFIG = (("CLASS_COHERENT", 0, 12, 0), ("CLASS_ANY", 0, 12, 0))
PMC_INCOHERENT = ((3, 8, 0), (3, 12, 0), (1, 12, 0))


def pmc_frame(pkg, bvh, step, dev, cnt):
    """Frames of the benchmark between two plain device copies of 512 MiB.
    The copies calibrate FETCH_SIZE / WRITE_SIZE (a copy of B bytes reads B
    and writes B); the frames give the per-dispatch counters of the traversal
    kernels with the launch configurations the benchmark times (two frames
    per incoherent candidate; each candidate is a different kernel)."""
    import torch
    for cls, sched, depth, bpc in PMC_FRAME_CONFIG:
        bvh.set_trace_config(getattr(pkg, cls), sched, depth, bpc)
    x = torch.ones(128 << 20, device=dev)
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    for sched, depth, bpc in PMC_INCOHERENT:
        bvh.set_trace_config(pkg.CLASS_INCOHERENT, sched, depth, bpc)
        for _ in range(2):
            step()
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    log(f"pmc frame: {bvh.num_bvh4_nodes} nodes, {int(cnt.item())} secondary rays, copy {x.numel() * 4} bytes")


def load_traffic(config):
    """HBM bytes per launch of the dominant kernel in launch configuration
    `config` ("schedule/lds_depth/blocks_per_cu"), measured with rocprofv3 PMC
    counters and committed under profiles/ (null when not measured)."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("incoherent_closest_bytes_per_launch", {}).get(config)


def build_report(bvh, tris, build_s):
    """The GPU build of the frame's BVH against its roofline: algorithmic bytes
    per triangle 36 (read) + 36 (reordered write) + 36 per binning pass (one
    per BVH2 level), plus 64 per BVH4 node, over the device build time."""
    st = bvh.build_stats()
    n = int(tris.shape[0])
    alg = n * (36 + 36 + 36 * st["levels"]) + 64 * bvh.num_bvh4_nodes
    dev_s = st["device_ms"] * 1e-3
    gbs = alg / dev_s / 1e9
    log(f"build: {build_s * 1e3:.1f} ms wall (upload {st['upload_ms']:.1f}, device {st['device_ms']:.1f}, "
        f"handle {st['handle_ms']:.1f}), {st['levels']} levels, {gbs:.0f} GB/s algorithmic")
    return {"wall_ms": round(build_s * 1e3, 2), "upload_ms": round(st["upload_ms"], 2),
            "device_ms": round(st["device_ms"], 2), "handle_ms": round(st["handle_ms"], 2),
            "levels": st["levels"], "bvh4_nodes": bvh.num_bvh4_nodes, "alg_bytes": int(alg),
            "achieved_GBps": round(gbs, 1), "peak_GBps": HBM_PEAK_GBS, "frac": round(gbs / HBM_PEAK_GBS, 4),
            "bound": "launch/latency (one pass per BVH2 level)"}


def cpu_build_baseline(bvh, tris):
    """The oracle's host build of the same triangles (one pass, one thread):
    its time, and whether its BVH4 nodes equal the GPU build's byte for byte."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    t0 = time.perf_counter()
    ob = pyoracle.OracleBVH(tris)
    cpu_s = time.perf_counter() - t0
    same = bool(np.array_equal(ob.nodes(), bvh.node_data()))
    log(f"cpu build (oracle, 1 thread): {cpu_s:.2f} s; GPU tree identical: {same}")
    return {"cpu_s": round(cpu_s, 3), "cpu_kind": "port", "cpu_threads": 1, "tree_identical_to_cpu": same}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(bvh, Pr, D, m, sample, budget_s=8.0, gpu_hits=None):
    """The CPU oracle's traversal (C, OpenMP over all host cores) on an evenly
    strided sample of the same primary and diffuse rays and the same BVH:
    watertight triangle test (the parity arithmetic, `value`) and the
    reference's Moller-Trumbore test (`value_mt`), each one warm-up pass then
    at least 5 timed passes (about budget_s seconds), median. gpu_hits
    ((primary ids, t), (diffuse ids, t)) device tensors: the watertight
    answers are compared with them ray by ray (`parity`)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    nodes = bvh.node_data()
    lt = np.ascontiguousarray(bvh.tri_data().T)
    pid = bvh.leaf_prim_ids()
    n = Pr["ox"].shape[0]
    ip = np.linspace(0, n - 1, min(sample, n)).astype(np.int64)
    idd = np.linspace(0, m - 1, min(sample, m)).astype(np.int64)
    prim = np.stack([Pr[k].cpu().numpy()[ip] for k in RAY_KEYS]).astype(np.float32)
    diff = np.stack([D[k][:m].cpu().numpy()[idd] for k in RAY_KEYS]).astype(np.float32)
    threads = pyoracle.num_threads(0)
    rays = prim.shape[1] + diff.shape[1]
    res = {}
    parity = None
    if gpu_hits is not None:
        (gi_p, gt_p), (gi_d, gt_d) = gpu_hits
        rp = pyoracle.trace_closest(nodes, lt, pid, prim, counts=False)
        rd = pyoracle.trace_closest(nodes, lt, pid, diff, counts=False)
        ids = np.concatenate([gi_p.cpu().numpy()[ip], gi_d[:m].cpu().numpy()[idd]])
        ts = np.concatenate([gt_p.cpu().numpy()[ip], gt_d[:m].cpu().numpy()[idd]]).view(np.uint32)
        rid = np.concatenate([rp["id"], rd["id"]])
        rt = np.concatenate([rp["t"], rd["t"]]).view(np.uint32)
        parity = {"rays_compared": int(rid.shape[0]), "id_mismatches": int((ids != rid).sum()),
                  "t_bit_mismatches": int((ts != rt).sum())}
        log(f"parity vs oracle: {parity}")
    for name, isect in (("watertight", 0), ("mt", 1)):
        pyoracle.trace_closest(nodes, lt, pid, prim, isect=isect, counts=False)       # warm-up pass
        times = []
        t_all = time.perf_counter()
        while len(times) < 5 or (time.perf_counter() - t_all < budget_s and len(times) < 50):
            t0 = time.perf_counter()
            pyoracle.trace_closest(nodes, lt, pid, prim, isect=isect, counts=False)
            pyoracle.trace_closest(nodes, lt, pid, diff, isect=isect, counts=False)
            times.append(time.perf_counter() - t0)
        res[name] = (float(np.median(times)), len(times))
        log(f"cpu baseline ({name}): {rays} rays in {res[name][0]:.3f} s (median of {len(times)}) on {threads} threads")
    tw, nw = res["watertight"]
    tm, nm = res["mt"]
    return {"value": round(rays / tw / 1e6, 2), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "parity": parity,
            "value_mt": round(rays / tm / 1e6, 2), "cpu_model": cpu_model(),
            "sample": f"{prim.shape[1]} primary + {diff.shape[1]} diffuse rays of the frame, closest hit, same BVH; "
                      f"value: watertight test (median of {nw} passes), value_mt: the reference's Moller-Trumbore "
                      f"test (median of {nm}); one warm-up pass each, ~{budget_s:.0f} s of CPU work each"}


if __name__ == "__main__":
    main()
