"""Benchmark of the path-tracing hot path on one MI355X.

One step = one full render of the workload (every sample of every pixel
through the wavefront bounce loop).  Default workload: the Cornell box with a
noise-displaced level-6 icosphere (81,934 primitives), 512x512, 8x8
sub-pixels x 1 sample = 64 spp, max depth 16, Russian roulette after depth 3,
FP32.  Prints one JSON line.

value = closest-hit segments per second (Mrays/s) over the timed steps, with
the scene resident in HBM before timing starts.  The roofline object prices
the dominant kernel (extend: BVH traversal) with its algorithmic bytes per
launch over its event-timed launch durations; the cpu_baseline object times
the CPU restatement (oracle/) on a band of rows of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "omochi-renderer3_amd"))

METRIC = "Mrays/s at matched RMSE, served/batched vs plain GPU; p50/p99 job latency"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes of the extend kernel: per segment the queue entry (4),
# ray origin+direction (32), job slot (4) and hit record (16); per node visit
# the six bound vectors and the child vector (112); per primitive test one
# 48-byte primitive record.
B_SEG, B_NODE, B_PRIM = 56, 112, 48

WORKLOADS = {
    "c0": dict(kind=0, w=256, h=256, ss=4, spp=1, depth=8, rr=3, env_is=0,
               desc="Cornell box (12 triangles + 2 spheres), 256x256, 16 spp, depth 8, RR after 3"),
    "c1": dict(kind=1, w=512, h=512, ss=8, spp=1, depth=16, rr=3, env_is=0,
               desc="Cornell box + displaced level-6 icosphere (81,934 primitives), 512x512, 64 spp, depth 16, "
                    "RR after 3"),
    "c2": dict(kind=2, w=1280, h=720, ss=16, spp=1, depth=12, rr=3, env_is=1,
               desc="heightfield terrain (997,002 triangles) + 200 spheres, 2048x1024 HDR sky with importance "
                    "sampling, 1280x720, 256 spp, depth 12"),
}


def cpu_baseline(scene, cam, prm, rows, threads):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    o = pyoracle.Oracle(scene)
    t0 = time.perf_counter()
    img, seg, sh = o.render(cam, prm, accel=pyoracle.OWN_BVH, threads=threads, rows=rows)
    dt = time.perf_counter() - t0
    o.close()
    return img, (seg + sh) / dt / 1e6, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c1", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-rows", type=int, default=64)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); multi-GPU sharding is design-only.",
              file=sys.stderr)
        sys.exit(2)
    import omr

    wl = WORKLOADS[args.workload]
    scene = omr.procedural(wl["kind"], width=wl["w"], height=wl["h"])
    ctx = omr.Context(0)
    dev = ctx.device_info()
    sid = ctx.upload(scene)
    info = ctx.scene_info(sid)
    prm = omr.params(ss=wl["ss"], spp_end=wl["spp"], max_depth=wl["depth"], rr_depth=wl["rr"], env_is=wl["env_is"],
                     seed=0x6F6D7233)
    ctx.set_profiling(True)
    for _ in range(args.warmup):
        ctx.render(sid, scene.camera, prm)
    seg = shadow = 0
    ext_ms = shade_ms = shadow_ms = 0.0
    launches = nodes = prims = 0
    img = None
    t0 = time.perf_counter()
    for _ in range(args.steps):
        img = ctx.render(sid, scene.camera, prm)  # returns after the device finished and copied back
        st = ctx.last_stats()
        seg += st["segments"]
        shadow += st["shadow_rays"]
        ext_ms += st["extend_ms"]
        shade_ms += st["shade_ms"]
        shadow_ms += st["shadow_ms"]
        launches += st["extend_launches"]
        nodes += st["node_visits"]
        prims += st["prim_tests"]
    dt = time.perf_counter() - t0
    rays = seg + shadow
    value = rays / dt / 1e6
    alg_bytes = B_SEG * seg + B_NODE * nodes + B_PRIM * prims
    achieved = alg_bytes / (ext_ms / 1e3) / 1e9 if ext_ms > 0 else 0.0
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"extend_{args.workload}_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    roofline = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
        "kernel": "k_extend", "avg_launch_ms": round(ext_ms / max(1, launches), 4),
        "alg_bytes_per_launch": int(alg_bytes / max(1, launches)),
        "nodes_per_segment": round(nodes / max(1, seg), 3), "prims_per_segment": round(prims / max(1, seg), 3),
        "kernel_ms_per_step": {"extend": round(ext_ms / args.steps, 3), "shade": round(shade_ms / args.steps, 3),
                               "shadow": round(shadow_ms / args.steps, 3)},
    }
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{args.workload}: {wl['desc']}", "mode": "plain (one job per render)",
                   "resolution": f"{wl['w']}x{wl['h']}", "spp": wl["ss"] ** 2 * wl["spp"],
                   "max_depth": wl["depth"], "rr_depth": wl["rr"], "env_is": wl["env_is"],
                   "primitives": scene.n_prims, "bvh_nodes": info["n_nodes"], "device": dev["name"],
                   "n_cu": dev["n_cu"]},
        "segments_per_step": seg // args.steps, "shadow_rays_per_step": shadow // args.steps,
        "roofline": roofline,
    }
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        r0 = max(0, wl["h"] // 2 - args.cpu_rows // 2)
        rows = (r0, r0 + args.cpu_rows)
        cimg, cmrays, cdt = cpu_baseline(scene, scene.camera, prm, rows, threads)
        diff = img[rows[0]:rows[1]].astype(np.float64) - cimg.astype(np.float64)
        out["rmse_vs_cpu"] = float(np.sqrt(np.mean(diff * diff)))
        out["cpu_baseline"] = {"value": round(cmrays, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
                               "sample": f"rows {rows[0]}-{rows[1] - 1} of the same render ({args.cpu_rows} x "
                                         f"{wl['w']} px, all {wl['ss'] ** 2 * wl['spp']} spp), {cdt:.1f} s wall"}
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
