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
# shade: queue index 4 + path state and hit 80 read, state 64 + queue 4 written, material 48 + shading-normal
# record 48 read per segment; a shadow record (48) per environment sample
B_SHADE_SEG, B_SHADOW_REC = 248, 48
# drain: queue index 4 + path state 68 read and the state word 4 written per path; material 48 + shading-normal
# record 48 per segment, plus the traversal's node and primitive bytes
B_DRAIN_PATH, B_DRAIN_SEG = 76, 96

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


def batched_jobs(n_jobs=64, seed=0x6F6D7233):
    """Config 3: uneven jobs over four resident scenes (Cornell, Cornell +
    icosphere, terrain + sky, torus knot).  Pixel count log-uniform in
    [64^2, 1024^2], spp log-uniform in [1, 256] (as ss^2 x samples), depth in
    {4, 8, 16}; environment importance sampling where a sky exists."""
    rng = np.random.default_rng(seed)
    jobs = []
    for j in range(n_jobs):
        kind = j % 4
        side = int(round(np.exp(rng.uniform(np.log(64), np.log(1024)))))
        spp = float(np.exp(rng.uniform(0.0, np.log(256))))
        ss = int(min(16, max(1, np.floor(np.sqrt(spp)))))
        k = max(1, int(round(spp / (ss * ss))))
        depth = int(rng.choice([4, 8, 16]))
        jobs.append(dict(kind=kind, w=side, h=side, ss=ss, spp=k, depth=depth, env_is=int(kind in (2, 3)),
                         seed=1000 + j))
    return jobs


# Config 4's eight resident scenes: (procedural kind, seed offset).  Kind 4 is
# a 2000x1000 terrain (3,996,002 triangles) + 200 spheres with its own sky.
SERVED_SCENES = [(0, 1), (1, 1), (1, 2), (2, 1), (3, 1), (3, 2), (4, 1), (0, 2)]


def served_requests(n, kinds, seed):
    """Config 4 requests: pixel count log-uniform in [64^2, 1920x1080] at 16:9,
    spp in {1, 4, 16, 64, 256} (ss = sqrt(spp)), random seeds."""
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(n):
        px = float(np.exp(rng.uniform(np.log(64 * 64), np.log(1920 * 1080))))
        w = max(16, int(round(np.sqrt(px * 16 / 9))))
        h = max(9, int(round(px / w)))
        ss = int(rng.choice([1, 2, 4, 8, 16]))
        si = int(rng.integers(0, len(kinds)))
        reqs.append(dict(scene=si, w=w, h=h, ss=ss, spp=1, depth=8, env_is=int(kinds[si][0] in (2, 3, 4)),
                         seed=int(rng.integers(1, 2 ** 31))))
    return reqs


def _prm(omr, j):
    return omr.params(ss=j["ss"], spp_end=j["spp"], max_depth=j["depth"], rr_depth=3, env_is=j["env_is"],
                      seed=j["seed"])


def run_batched(ctx, omr, args):
    kinds = [0, 1, 2, 3]
    scenes = {k: omr.procedural(k, width=64, height=64) for k in kinds}
    sids = {k: ctx.upload(s) for k, s in scenes.items()}
    jobs = batched_jobs()
    cams = [omr.procedural_camera(j["kind"], j["w"], j["h"]) for j in jobs]
    prms = [_prm(omr, j) for j in jobs]
    ids = [sids[j["kind"]] for j in jobs]
    tune = None
    if not args.no_tune:
        # tune on the workload itself: a subset of the jobs favours smaller
        # pools than the whole batch wants
        tune = ctx.autotune(ids, cams, prms, budget_ms=args.tune_ms)
    # plain: each job alone, one after another
    t0 = time.perf_counter()
    plain, rays_plain = [], 0
    for i in range(len(jobs)):
        plain.append(ctx.render(ids[i], cams[i], prms[i]))
        st = ctx.last_stats()
        rays_plain += st["segments"] + st["shadow_rays"]
    t_plain = time.perf_counter() - t0
    # batched: all jobs in one persistent wavefront run
    for _ in range(args.warmup):
        ctx.render_batch(ids, cams, prms)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch = ctx.render_batch(ids, cams, prms)
    t_batch = (time.perf_counter() - t0) / args.steps
    st = ctx.last_stats()
    rays_batch = st["segments"] + st["shadow_rays"]
    identical = all(np.array_equal(a, b) for a, b in zip(plain, batch))
    return dict(jobs=jobs, plain_mrays=rays_plain / t_plain / 1e6, batch_mrays=rays_batch / t_batch / 1e6,
                t_plain=t_plain, t_batch=t_batch, rays=rays_batch, identical=identical, tuning=tune,
                n_prims=sum(s.n_prims for s in scenes.values()))


def run_served(ctx, omr, args):
    kinds = SERVED_SCENES
    scenes = [omr.procedural(k, seed=0x6F6D7233 + s, width=64, height=64) for k, s in kinds]
    sids = [ctx.upload(s) for s in scenes]
    n_tri = sum(s.tri.shape[0] for s in scenes)
    cal = served_requests(args.requests, kinds, seed=7)

    def job_args(r):
        cam = omr.procedural_camera(kinds[r["scene"]][0], r["w"], r["h"])
        return sids[r["scene"]], cam, _prm(omr, r)

    # plain: each request alone, back to back (also gives the service rate)
    t0 = time.perf_counter()
    rays_plain = 0
    alone = []
    for r in cal:
        alone.append(ctx.render(*job_args(r)))
        st = ctx.last_stats()
        rays_plain += st["segments"] + st["shadow_rays"]
    t_plain = time.perf_counter() - t0
    capacity = len(cal) / t_plain
    rates = []
    rng = np.random.default_rng(11)
    for f in (0.5, 0.8, 1.1):
        lam = f * capacity
        ctx.server_start()
        outs = [np.zeros((r["h"], r["w"], 3), np.float32) for r in cal]
        arrivals = np.cumsum(rng.exponential(1.0 / lam, len(cal)))
        t_start = time.perf_counter()
        ids = []
        for r, o, a in zip(cal, outs, arrivals):
            now = time.perf_counter() - t_start
            if a > now:
                time.sleep(a - now)
            ids.append(ctx.submit(*job_args(r), o))
        for i in ids:
            if ctx.wait(i, 600000) != 0:
                raise RuntimeError("served job failed")
        t_all = time.perf_counter() - t_start
        st = ctx.server_stats()
        ctx.server_stop()
        same = all(np.array_equal(a, b) for a, b in zip(alone, outs))
        rays = st["segments"] + st["shadow_rays"]
        rates.append(dict(load=f, rate_jobs_per_s=round(lam, 3), p50_ms=round(st["lat_p50_ms"], 2),
                          p99_ms=round(st["lat_p99_ms"], 2), mrays=round(rays / st["busy_ms"] / 1e3, 2),
                          wall_s=round(t_all, 2), identical_to_alone=same))
    return dict(plain_mrays=rays_plain / t_plain / 1e6, capacity=capacity, rates=rates, n_tri=n_tri,
                n_requests=len(cal))


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
    ap.add_argument("--steps", type=int, default=None,
                    help="timed renders (default: 20 for the plain workloads, 3 batches for c3)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed renders first (default 2; 1 for c3)")
    ap.add_argument("--workload", default="c1", choices=sorted(WORKLOADS) + ["c3", "c4"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-rows", type=int, default=0, help="rows of the CPU baseline sample (0 = all)")
    ap.add_argument("--no-tune", action="store_true", help="c3: skip the autotuner")
    ap.add_argument("--tune-ms", type=int, default=20000)
    ap.add_argument("--requests", type=int, default=32, help="c4: requests per arrival rate")
    ap.add_argument("--no-profile", action="store_true",
                    help="time the renders without per-launch events (no roofline timing)")
    ap.add_argument("--tune", default="", help="launch tuning overrides, e.g. wg_size=64,blocks_per_cu=16")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="after the run, copy MIB MiB device-to-device three times (hardware-counter calibration)")
    args = ap.parse_args()
    plain = args.workload not in ("c3", "c4")
    if args.steps is None:
        args.steps = 20 if plain else 3
    if args.warmup is None:
        args.warmup = 2 if plain else 1
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); multi-GPU sharding is design-only.",
              file=sys.stderr)
        sys.exit(2)
    import omr

    if args.workload in ("c3", "c4"):
        ctx = omr.Context(0)
        dev = ctx.device_info()
        if args.workload == "c3":
            r = run_batched(ctx, omr, args)
            out = {
                "metric": METRIC, "value": round(r["batch_mrays"], 2), "unit": "Mrays/s", "n_gpus": 1,
                "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(r["t_batch"] * 1e3, 3),
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
                "data": "synthetic",
                "config": {"workload": "c3: 64 uneven jobs over 4 resident scenes (Cornell, Cornell + icosphere, "
                                       "terrain + sky, torus knot), one persistent wavefront run",
                           "mode": "batched", "primitives": r["n_prims"], "device": dev["name"]},
                "plain_mrays": round(r["plain_mrays"], 2), "plain_ms": round(r["t_plain"] * 1e3, 1),
                "speedup_vs_plain": round(r["batch_mrays"] / r["plain_mrays"], 3),
                "images_identical_to_plain": r["identical"], "rays_per_step": r["rays"], "tuning": r["tuning"],
            }
        else:
            r = run_served(ctx, omr, args)
            mid = r["rates"][1]
            out = {
                "metric": METRIC, "value": mid["mrays"], "unit": "Mrays/s", "n_gpus": 1, "steps": 1, "warmup": 0,
                "ms_per_step": round(mid["wall_s"] * 1e3, 1), "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None, "dtype": "f32", "data": "synthetic",
                "config": {"workload": f"c4: Poisson arrivals over 8 resident scenes ({r['n_tri']} triangles), "
                                       f"{r['n_requests']} requests per rate", "mode": "served",
                           "device": dev["name"]},
                "plain_mrays": round(r["plain_mrays"], 2), "capacity_jobs_per_s": round(r["capacity"], 3),
                "rates": r["rates"], "p50_ms": mid["p50_ms"], "p99_ms": mid["p99_ms"],
            }
        print(json.dumps(out), flush=True)
        ctx.close()
        return

    wl = WORKLOADS[args.workload]
    scene = omr.procedural(wl["kind"], width=wl["w"], height=wl["h"])
    ctx = omr.Context(0)
    dev = ctx.device_info()
    sid = ctx.upload(scene)
    info = ctx.scene_info(sid)
    if args.tune:
        ctx.set_tuning(**{k: int(v) for k, v in (kv.split("=") for kv in args.tune.split(","))})
    prm = omr.params(ss=wl["ss"], spp_end=wl["spp"], max_depth=wl["depth"], rr_depth=wl["rr"], env_is=wl["env_is"],
                     seed=0x6F6D7233)
    ctx.set_profiling(not args.no_profile)
    for _ in range(args.warmup):
        ctx.render(sid, scene.camera, prm)
    seg = shadow = drain_seg = drained = 0
    ext_ms = shade_ms = shadow_ms = drain_ms = 0.0
    launches = nodes = prims = dnodes = dprims = 0
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
        drain_seg += st["drain_segments"]
        drained += st["drained_paths"]
        drain_ms += st["drain_ms"]
        dnodes += st["drain_node_visits"]
        dprims += st["drain_prim_tests"]
    dt = time.perf_counter() - t0
    rays = seg + shadow
    value = rays / dt / 1e6
    ext_seg = seg - drain_seg  # segments of the extend launches
    drains = args.steps if drained else 0  # one drain launch per render
    # algorithmic bytes of each kernel over the timed renders (per-unit figures in DESIGN.md section 3)
    alg = {
        "k_extend": (B_SEG * ext_seg + B_NODE * nodes + B_PRIM * prims, ext_ms, launches),
        "k_shade": (B_SHADE_SEG * ext_seg + B_SHADOW_REC * shadow, shade_ms, launches),
        "k_drain": (B_DRAIN_PATH * drained + B_DRAIN_SEG * drain_seg + B_NODE * dnodes + B_PRIM * dprims, drain_ms,
                    drains),
    }
    kernels = {}
    for name, (b, ms, n) in alg.items():
        if ms > 0 and n > 0:
            a = b / (ms / 1e3) / 1e9
            kernels[name] = {"ms_per_step": round(ms / args.steps, 3), "avg_launch_ms": round(ms / n, 4),
                             "alg_bytes_per_launch": int(b / n), "achieved": round(a, 1),
                             "frac": round(a / HBM_PEAK_GBS, 4)}
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_step"]) if kernels else "k_extend"
    kd = kernels.get(dom, {"achieved": 0.0, "frac": 0.0, "avg_launch_ms": 0.0, "alg_bytes_per_launch": 0})
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"{dom[2:]}_{args.workload}_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    roofline = {
        "bound": "hbm", "achieved": kd["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": kd["frac"], "traffic": traffic,
        "kernel": dom, "avg_launch_ms": kd["avg_launch_ms"], "alg_bytes_per_launch": kd["alg_bytes_per_launch"],
        "kernels": kernels,
        "nodes_per_segment": round((nodes + dnodes) / max(1, seg), 3),
        "prims_per_segment": round((prims + dprims) / max(1, seg), 3),
        "kernel_ms_per_step": {"extend": round(ext_ms / args.steps, 3), "shade": round(shade_ms / args.steps, 3),
                               "shadow": round(shadow_ms / args.steps, 3), "drain": round(drain_ms / args.steps, 3)},
        "drain": {"paths_per_step": drained // args.steps, "segments_per_step": drain_seg // args.steps},
    }
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{args.workload}: {wl['desc']}", "mode": "plain (one job per render)",
                   "resolution": f"{wl['w']}x{wl['h']}", "spp": wl["ss"] ** 2 * wl["spp"],
                   "max_depth": wl["depth"], "rr_depth": wl["rr"], "env_is": wl["env_is"],
                   "primitives": scene.n_prims, "bvh_nodes": info["n_nodes"], "device": dev["name"],
                   "n_cu": dev["n_cu"], "tuning": ctx.get_tuning()},
        "segments_per_step": seg // args.steps, "shadow_rays_per_step": shadow // args.steps,
        "roofline": roofline,
    }
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        nrows = args.cpu_rows if 0 < args.cpu_rows < wl["h"] else wl["h"]
        r0 = max(0, wl["h"] // 2 - nrows // 2)
        rows = (r0, r0 + nrows)
        cimg, cmrays, cdt = cpu_baseline(scene, scene.camera, prm, rows, threads)
        diff = img[rows[0]:rows[1]].astype(np.float64) - cimg.astype(np.float64)
        out["rmse_vs_cpu"] = float(np.sqrt(np.mean(diff * diff)))
        out["cpu_baseline"] = {"value": round(cmrays, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
                               "sample": f"rows {rows[0]}-{rows[1] - 1} of the same render ({nrows} x "
                                         f"{wl['w']} px, all {wl['ss'] ** 2 * wl['spp']} spp), {cdt:.1f} s wall"}
    print(json.dumps(out), flush=True)
    ctx.close()
    if args.calib_copy > 0:
        copy_calibration(args.calib_copy << 20)


def copy_calibration(nbytes):
    """Known-size traffic for scaling hardware byte counters: one fill of
    nbytes and three device-to-device copies of nbytes (each reads and writes
    nbytes once)."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    if hip.hipMemset(src, 1, C.c_size_t(nbytes)) != 0:
        raise RuntimeError("hipMemset failed")
    for _ in range(3):
        if hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_calibration_bytes": nbytes, "copies": 3}), flush=True)


if __name__ == "__main__":
    main()
