"""Benchmark of the path-tracing hot path on one MI355X.  Prints one JSON line.

Default workload (the configuration the headline metric is quoted on): c4,
served rendering.  Eight procedural scenes (5.75M triangles, two HDR skies)
stay resident; a seeded generator draws render requests (pixel count
log-uniform in 64^2..1920x1080, spp in {1,4,16,64,256}, depth 8, random
seeds).  One step is one Poisson stream of --requests requests offered at
--load (default 2.0) times the plain path's capacity, so the server, not the
arrival process, bounds the rate; served by continuous batching (requests
admitted while others run, rays of all live requests in shared queues).

  value            closest-hit segments + shadow rays per second over the K
                   timed streams (first submission to last result on the
                   host), scenes resident before timing starts
  plain_mrays      the same requests rendered one after another, each alone
                   (omr_render), wall clock
  latency          p50 / p99 submission-to-result latency at 0.5 and 0.8 of
                   the plain capacity and at --load (over all K streams)
  batched          config 3 in the same run: 64 uneven jobs over four scenes
                   in one persistent launch vs the same jobs alone
  by_size          where packing wins: rays and plain Mrays/s per request size
  roofline         the dominant kernel of a profiled plain pass over the first
                   --roofline-requests requests: algorithmic bytes (the
                   survey's per-unit figures) over its event-timed launches,
                   measured HBM traffic (rocprof, scaled by a copy
                   calibration), and the VALU issue share that bounds it
  cpu_baseline     the CPU restatement of the path (oracle/) on all usable
                   host cores over a bounded sample of the same requests

--workload c0|c1|c2 measure one plain render configuration (20 timed renders);
--workload c3 the batched configuration alone.
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
# Algorithmic bytes per unit (the survey's roofline model): a closest-hit
# segment moves its ray in (32) and hit out (16) in the traversal kernel and
# the path state in and out (96) in shading; a node visit reads its node (128
# bytes 4-wide, 256 bytes 8-wide); a primitive test reads 48 (a sphere's
# 16-byte record is priced as a triangle: sphere tests are < 0.1 % of the
# tests here).  The drain does both halves.
B_SEG_TRACE, B_SEG_SHADE, B_PRIM = 48, 96, 48
B_NODE = {4: 128, 8: 256}
B_SHADOW = 48  # a shadow ray: its record in and the visibility result
# MI355X: 256 CUs x 4 SIMDs, one wave64 VALU instruction per SIMD every 4
# cycles at 2.4 GHz
VALU_ISSUE_PER_S = 256 * 4 * 2.4e9 / 4

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

# Config 4's eight resident scenes: (procedural kind, seed offset).  Kind 4 is
# a 2000x1000 terrain (3,996,002 triangles) + 200 spheres with its own sky.
SERVED_SCENES = [(0, 1), (1, 1), (1, 2), (2, 1), (3, 1), (3, 2), (4, 1), (0, 2)]
SIZE_BUCKETS = [0, 1 << 20, 1 << 23, 1 << 26, 1 << 40]  # samples per job


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


def log(msg):
    """Progress on stderr (stdout carries only the JSON line)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def _prm(omr, j):
    return omr.params(ss=j["ss"], spp_end=j["spp"], max_depth=j["depth"], rr_depth=3, env_is=j["env_is"],
                      seed=j["seed"])


def _samples(j):
    return j["w"] * j["h"] * j["ss"] * j["ss"] * j["spp"]


def by_size(jobs, rays, wall_s, extra=None):
    """Per request-size bucket: count, share of all rays, and Mrays/s of the
    jobs rendered alone (small jobs cannot fill the GPU alone: that is where
    packing them with others pays)."""
    out = []
    tot = float(sum(rays))
    for lo, hi in zip(SIZE_BUCKETS[:-1], SIZE_BUCKETS[1:]):
        idx = [i for i, j in enumerate(jobs) if lo <= _samples(j) < hi]
        if not idx:
            continue
        r = sum(rays[i] for i in idx)
        t = sum(wall_s[i] for i in idx)
        b = {"samples": f"[{lo}, {hi})" if hi < SIZE_BUCKETS[-1] else f">= {lo}", "jobs": len(idx),
             "share_of_rays": round(r / tot, 4), "alone_mrays": round(r / t / 1e6, 1),
             "alone_ms_mean": round(t / len(idx) * 1e3, 3)}
        if extra:
            b.update(extra(idx))
        out.append(b)
    return out


def render_alone(ctx, args_list):
    """Each job alone, back to back: images, rays and wall seconds per job."""
    imgs, rays, walls = [], [], []
    for a in args_list:
        t0 = time.perf_counter()
        imgs.append(ctx.render(*a))
        walls.append(time.perf_counter() - t0)
        st = ctx.last_stats()
        rays.append(st["segments"] + st["shadow_rays"])
    return imgs, rays, walls


def serve_stream(ctx, reqs, job_args, lam, rng, alone):
    """One Poisson stream at rate lam (jobs/s): returns rays, wall seconds
    (first submission to last result), per-request latencies (ms) and
    whether every result equals the request rendered alone."""
    ctx.server_start()
    try:
        outs = [np.empty((r["h"], r["w"], 3), np.float32) for r in reqs]
        arrivals = np.cumsum(rng.exponential(1.0 / lam, len(reqs)))
        arrivals -= arrivals[0]
        ids = []
        t0 = time.perf_counter()
        for r, o, a in zip(reqs, outs, arrivals):
            now = time.perf_counter() - t0
            if a > now:
                time.sleep(a - now)
            ids.append(ctx.submit(*job_args(r), o))
        for i in ids:
            rc = ctx.wait(i, 600000)
            if rc != 0:
                raise RuntimeError(f"served request failed ({rc}): {ctx.job_error(i)}")
        wall = time.perf_counter() - t0
        lat = [ctx.job_latency(i) for i in ids]
        st = ctx.server_stats()
    finally:
        ctx.server_stop()
    same = all(np.array_equal(a, b) for a, b in zip(alone, outs))
    return st["segments"] + st["shadow_rays"], wall, lat, same


def pct(v, q):
    v = sorted(v)
    return float(v[min(len(v) - 1, int(round(q * (len(v) - 1))))])


# ---------------------------------------------------------------- roofline
def profiled_pass(ctx, args_list):
    """Renders with per-launch HIP events on the engine's stream (profiling
    on: every bounce-step kernel launched directly, node visits and primitive
    tests counted by the extend and drain kernels' counting variants)."""
    keys = ("segments", "shadow_rays", "extend_ms", "shade_ms", "shadow_ms", "extend_launches", "shadow_launches",
            "node_visits",
            "prim_tests", "drain_segments", "drained_paths", "drain_ms", "drain_node_visits", "drain_prim_tests")
    acc = dict.fromkeys(keys, 0)
    drains = 0
    ctx.set_profiling(True)
    try:
        for a in args_list:
            ctx.render(*a)
            st = ctx.last_stats()
            for k in keys:
                acc[k] += st[k]
            drains += st["drain_launches"]
    finally:
        ctx.set_profiling(False)
    acc["drain_launches"] = drains
    return acc


def roofline_from(acc, workload, node_width=4):
    ext_seg = acc["segments"] - acc["drain_segments"]
    bn, w = B_NODE[node_width], node_width
    kernels = {
        # name: (algorithmic bytes, summed event ms, launches, rocprof symbol of the counting variant)
        "k_extend": (B_SEG_TRACE * ext_seg + bn * acc["node_visits"] + B_PRIM * acc["prim_tests"],
                     acc["extend_ms"], acc["extend_launches"], f"k_extend<64, 16, {w}, true, true, true>"),
        # (single-job shading and drain kernels come in two forms, for scenes
        # with and without an environment map: the last template argument)
        "k_shade": (B_SEG_SHADE * ext_seg, acc["shade_ms"], acc["extend_launches"], "k_shade<64, true, true, *>"),
        "k_shadow": (B_SHADOW * acc["shadow_rays"], acc["shadow_ms"], acc["shadow_launches"],
                     f"k_shadow<64, 16, {w}, true, true, true>"),
        "k_drain": ((B_SEG_TRACE + B_SEG_SHADE) * acc["drain_segments"] + bn * acc["drain_node_visits"]
                    + B_PRIM * acc["drain_prim_tests"], acc["drain_ms"], acc["drain_launches"],
                    f"k_drain<64, 16, {w}, true, true, *>"),
    }
    table = {}
    for name, (b, ms, n, sym) in kernels.items():
        if ms > 0 and n > 0:
            a = b / (ms / 1e3) / 1e9
            table[name] = {"ms_total": round(ms, 3), "launches": n, "avg_launch_ms": round(ms / n, 4),
                           "alg_bytes_per_launch": int(b / n), "achieved_GBs": round(a, 1),
                           "frac": round(a / HBM_PEAK_GBS, 4), "symbol": sym}
    dom = max(table, key=lambda k: table[k]["ms_total"])
    kd = table[dom]
    # measured HBM traffic and VALU issue of the same kernel (rocprof counters
    # of the same pass, committed under profiles/)
    traffic, valu = None, None
    tpath = os.path.join(ROOT, "profiles", f"{dom[2:]}_{workload}_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    vpath = os.path.join(ROOT, "profiles", f"{dom[2:]}_{workload}_valu.json")
    if os.path.exists(vpath):
        with open(vpath) as f:
            v = json.load(f)
        insts = v.get("valu_insts_per_launch")
        if insts:
            valu = {"valu_insts_per_launch": insts,
                    "issue_frac": round(insts / (kd["avg_launch_ms"] / 1e3) / VALU_ISSUE_PER_S, 4),
                    "peak_wave_insts_per_s": VALU_ISSUE_PER_S, "source": os.path.relpath(vpath, ROOT)}
    roof = {"bound": "hbm", "achieved": kd["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": kd["frac"], "traffic": traffic, "kernel": dom, "symbol": kd["symbol"],
            "avg_launch_ms": kd["avg_launch_ms"], "alg_bytes_per_launch": kd["alg_bytes_per_launch"]}
    if traffic:
        hbm = traffic / (kd["avg_launch_ms"] / 1e3) / 1e9
        roof["hbm_measured_GBs"] = round(hbm, 1)
        roof["hbm_measured_frac"] = round(hbm / HBM_PEAK_GBS, 4)
    if valu:
        roof["valu_issue"] = valu
        # what actually binds the kernel: the node and primitive bytes come from
        # L2 and the MALL (measured HBM traffic well under the algorithmic
        # bytes), and the VALU issue slots are mostly full
        roof["binding"] = {"limit": "VALU issue", "issue_frac": valu["issue_frac"],
                           "hbm_measured_frac": roof.get("hbm_measured_frac")}
    roof["kernels"] = table
    seg = max(1, acc["segments"])
    roof["nodes_per_segment"] = round((acc["node_visits"] + acc["drain_node_visits"]) / seg, 3)
    roof["prims_per_segment"] = round((acc["prim_tests"] + acc["drain_prim_tests"]) / seg, 3)
    return roof


# ---------------------------------------------------------------- CPU baseline
def host_cpu():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return model, os.cpu_count() or 1, usable


def cpu_threads():
    # the GPU host gives one GPU's job a share of 16 cores, whatever the
    # machine has in all
    _, _, usable = host_cpu()
    return max(1, min(16, usable))


def cpu_sample(items, seconds, threads):
    """items: (scene, camera, params, gpu image) per request.  Renders bands
    of 8 rows of each request in turn on the CPU restatement until `seconds`
    of wall time are spent; returns Mrays/s, the sample, and the RMSE of the
    CPU rows against the GPU rows."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    oracles, rays, t_cpu, sq, n_px = {}, 0, 0.0, 0.0, 0
    cursor = [0] * len(items)
    done_rows = 0
    used = set()
    while t_cpu < seconds:
        progressed = False
        for k, (scene, cam, prm, gimg) in enumerate(items):
            if t_cpu >= seconds or cursor[k] >= cam.height:
                continue
            if id(scene) not in oracles:
                oracles[id(scene)] = pyoracle.Oracle(scene)  # preparation (own BVH build) is not timed
            r0, r1 = cursor[k], min(cam.height, cursor[k] + 8)
            t0 = time.perf_counter()
            img, seg, sh = oracles[id(scene)].render(cam, prm, accel=pyoracle.OWN_BVH, threads=threads, rows=(r0, r1))
            t_cpu += time.perf_counter() - t0
            rays += seg + sh
            d = img.astype(np.float64) - gimg[r0:r1].astype(np.float64)
            sq += float((d * d).sum())
            n_px += d.size
            cursor[k] = r1
            done_rows += r1 - r0
            used.add(k)
            progressed = True
        if not progressed:
            break
    for o in oracles.values():
        o.close()
    return rays / t_cpu / 1e6, t_cpu, done_rows, len(used), float(np.sqrt(sq / max(1, n_px)))


def cpu_calibration():
    """The restatement's speed over the reference's compiled PathTracer on
    one thread (oracle/cpu_calibration.json, written where the reference
    checkout is present): the geometric mean over its workloads, the per
    workload ratios and the CPU they were measured on."""
    path = os.path.join(ROOT, "oracle", "cpu_calibration.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        doc = json.load(f)
    per = {r["workload"]: r["restatement_over_reference_bvh"]["t1"] for r in doc["results"]}
    gm = float(np.exp(np.mean(np.log(list(per.values())))))
    return {"restatement_over_reference": round(gm, 3), "per_workload_1_thread": per, "measured_on": doc["cpu_model"],
            "source": "oracle/cpu_calibration.json (oracle/calibrate_cpu.py: the reference's PathTracer compiled "
                      "from its sources, binary BVH, vs the restatement, same scenes, cameras and sample counts)"}


# ---------------------------------------------------------------- workloads
def run_batched(ctx, omr, args):
    kinds = [0, 1, 2, 3]
    scenes = {k: omr.procedural(k, width=64, height=64) for k in kinds}
    sids = {k: ctx.upload(s) for k, s in scenes.items()}
    jobs = batched_jobs()
    cams = [omr.procedural_camera(j["kind"], j["w"], j["h"]) for j in jobs]
    prms = [_prm(omr, j) for j in jobs]
    ids = [sids[j["kind"]] for j in jobs]
    alist = list(zip(ids, cams, prms))
    saved = ctx.get_tuning()
    tune = None
    log("c3: scenes resident")
    if args.tune_ms > 0:
        # the autotuner picks launch shape, pool, cadence and stack depth on
        # the batch itself; the same settings then serve both forms
        tune = ctx.autotune(ids, cams, prms, budget_ms=args.tune_ms)
    log(f"c3: tuning {tune}")
    render_alone(ctx, alist[:4])  # warm-up
    plain, rays_alone, walls = render_alone(ctx, alist)
    t_plain = sum(walls)
    ctx.render_batch(ids, cams, prms)  # warm-up
    t_batch, reps = 0.0, max(1, args.batch_reps)
    for _ in range(reps):
        t0 = time.perf_counter()
        batch = ctx.render_batch(ids, cams, prms)
        t_batch += time.perf_counter() - t0
    t_batch /= reps
    st = ctx.last_stats()
    rays_batch = st["segments"] + st["shadow_rays"]
    identical = all(np.array_equal(a, b) for a, b in zip(plain, batch))
    ctx.set_tuning(**saved)
    for s in sids.values():
        ctx.release(s)
    return {
        "workload": "c3: 64 uneven jobs over 4 resident scenes (Cornell, Cornell + icosphere, terrain + sky, "
                    "torus knot), one persistent wavefront run",
        "plain_mrays": round(sum(rays_alone) / t_plain / 1e6, 2), "batch_mrays": round(rays_batch / t_batch / 1e6, 2),
        "speedup_vs_plain": round((rays_batch / t_batch) / (sum(rays_alone) / t_plain), 3),
        "plain_ms": round(t_plain * 1e3, 1), "batch_ms": round(t_batch * 1e3, 1), "batch_reps": reps,
        "rays": rays_batch, "rays_equal_alone": rays_batch == sum(rays_alone), "identical_to_alone": identical,
        "tuning": tune, "primitives": sum(s.n_prims for s in scenes.values()),
        "by_size": by_size(jobs, rays_alone, walls),
    }


def run_served(ctx, omr, args, dev):
    kinds = SERVED_SCENES
    log("c4: building the eight scenes")
    scenes = [omr.procedural(k, seed=0x6F6D7233 + s, width=64, height=64) for k, s in kinds]
    sids = [ctx.upload(s) for s in scenes]
    n_tri = sum(s.tri.shape[0] for s in scenes)
    reqs = served_requests(args.requests, kinds, seed=7)

    def job_args(r):
        return sids[r["scene"]], omr.procedural_camera(kinds[r["scene"]][0], r["w"], r["h"]), _prm(omr, r)

    alist = [job_args(r) for r in reqs]
    # warm-up: a few requests alone and one short served stream (allocations,
    # graph captures)
    render_alone(ctx, alist[:max(1, args.warmup)])
    if args.roofline_only:
        # only the profiled pass below (one host thread: for hardware-counter
        # and kernel-trace runs of the roofline kernel)
        nroof = min(len(reqs), args.roofline_requests)
        log(f"c4: profiled pass over {nroof} requests")
        roof = roofline_from(profiled_pass(ctx, alist[:nroof]), "c4", ctx.get_tuning()["node_width"])
        return {"workload": "c4", "roofline_only": True, "roofline": roof}, []
    rng = np.random.default_rng(11)
    # plain: each request alone, back to back; also the service rate
    log(f"c4: {len(reqs)} requests alone")
    alone, rays_alone, walls = render_alone(ctx, alist)
    t_plain = sum(walls)
    rays_plain = sum(rays_alone)
    capacity = len(reqs) / t_plain
    log(f"c4: plain {rays_plain / t_plain / 1e6:.1f} Mrays/s, {t_plain:.2f} s")
    serve_stream(ctx, reqs[:16], job_args, 4 * capacity, rng, alone[:16])
    # timed: K streams offered at `load` x the plain capacity
    lam = args.load * capacity
    rays_s, wall_s, lat_s, same_all = 0, 0.0, [], True
    for k in range(args.steps):
        rays, wall, lat, same = serve_stream(ctx, reqs, job_args, lam, rng, alone)
        log(f"c4: served stream {k + 1}/{args.steps}: {rays / wall / 1e6:.1f} Mrays/s, identical {same}")
        rays_s += rays
        wall_s += wall
        lat_s += lat
        same_all = same_all and same and rays == rays_plain
    served_mrays = rays_s / wall_s / 1e6
    latency = []
    for f in (0.5, 0.8):
        rays, wall, lat, same = serve_stream(ctx, reqs, job_args, f * capacity, rng, alone)
        log(f"c4: load {f}: p50 {pct(lat, 0.5):.1f} ms, p99 {pct(lat, 0.99):.1f} ms")
        same_all = same_all and same and rays == rays_plain
        latency.append({"load": f, "rate_jobs_per_s": round(f * capacity, 3), "requests": len(lat),
                        "p50_ms": round(pct(lat, 0.5), 2), "p99_ms": round(pct(lat, 0.99), 2),
                        "max_ms": round(max(lat), 2), "mrays": round(rays / wall / 1e6, 1)})
    latency.append({"load": args.load, "rate_jobs_per_s": round(lam, 3), "requests": len(lat_s),
                    "p50_ms": round(pct(lat_s, 0.5), 2), "p99_ms": round(pct(lat_s, 0.99), 2),
                    "max_ms": round(max(lat_s), 2), "mrays": round(served_mrays, 1)})
    # latency per size bucket in the K timed streams
    lat_by_req = np.array(lat_s).reshape(args.steps, len(reqs))

    def extra(idx):
        l = lat_by_req[:, idx].ravel()
        return {"served_p50_ms": round(float(np.median(l)), 2)}

    sizes = by_size(reqs, rays_alone, walls, extra)
    # roofline: a profiled plain pass over the first requests
    nroof = min(len(reqs), args.roofline_requests)
    log(f"c4: profiled pass over {nroof} requests")
    acc = profiled_pass(ctx, alist[:nroof])
    roof = roofline_from(acc, "c4", ctx.get_tuning()["node_width"])
    roof["pass"] = (f"the first {nroof} requests rendered alone with per-launch HIP events on the engine's "
                    "stream (counting kernel variants)")
    out = {
        "metric": METRIC, "value": round(served_mrays, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall_s / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"c4: served, Poisson arrivals offered at {args.load}x plain capacity, {len(reqs)} "
                               "requests per step "
                               f"over 8 resident scenes ({n_tri} triangles, 2 HDR skies); pixels log-uniform in "
                               "64^2..1920x1080, spp in {1,4,16,64,256}, depth 8",
                   "mode": "served (continuous batching, HBM-budget admission)", "requests_per_step": len(reqs),
                   "rays_per_step": rays_plain, "device": dev["name"], "n_cu": dev["n_cu"]},
        "plain_mrays": round(rays_plain / t_plain / 1e6, 2), "plain_ms": round(t_plain * 1e3, 1),
        "speedup_vs_plain": round(served_mrays / (rays_plain / t_plain / 1e6), 3),
        "identical_to_alone": same_all, "capacity_jobs_per_s": round(capacity, 3),
        "p50_ms": latency[-1]["p50_ms"], "p99_ms": latency[-1]["p99_ms"], "latency": latency,
        "by_size": sizes, "roofline": roof,
    }
    items = [(scenes[r["scene"]], alist[i][1], alist[i][2], alone[i]) for i, r in enumerate(reqs[:8])]
    return out, items


def run_plain(ctx, omr, args, dev):
    wl = WORKLOADS[args.workload]
    scene = omr.procedural(wl["kind"], width=wl["w"], height=wl["h"])
    sid = ctx.upload(scene)
    info = ctx.scene_info(sid)
    prm = omr.params(ss=wl["ss"], spp_end=wl["spp"], max_depth=wl["depth"], rr_depth=wl["rr"], env_is=wl["env_is"],
                     seed=0x6F6D7233)
    a = (sid, scene.camera, prm)
    for _ in range(args.warmup):
        ctx.render(*a)
    imgs, rays, walls = render_alone(ctx, [a] * args.steps)
    dt = sum(walls)
    acc = profiled_pass(ctx, [a] * min(args.steps, 5))
    roof = roofline_from(acc, args.workload, ctx.get_tuning()["node_width"])
    roof["pass"] = f"{min(args.steps, 5)} renders with per-launch HIP events on the engine's stream"
    out = {
        "metric": METRIC, "value": round(sum(rays) / dt / 1e6, 2), "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{args.workload}: {wl['desc']}", "mode": "plain (one job per render)",
                   "resolution": f"{wl['w']}x{wl['h']}", "spp": wl["ss"] ** 2 * wl["spp"], "max_depth": wl["depth"],
                   "rr_depth": wl["rr"], "env_is": wl["env_is"], "primitives": scene.n_prims,
                   "bvh_nodes": info["n_nodes"], "device": dev["name"], "n_cu": dev["n_cu"],
                   "tuning": ctx.get_tuning()},
        "rays_per_step": rays[0], "roofline": roof,
    }
    return out, [(scene, scene.camera, prm, imgs[-1])]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps: served streams (c4, default 20) or renders (c0-c2, default 20)")
    ap.add_argument("--warmup", type=int, default=5, help="untimed renders first")
    ap.add_argument("--workload", default="c4", choices=sorted(WORKLOADS) + ["c3", "c4"])
    ap.add_argument("--requests", type=int, default=128, help="c4: requests per served stream")
    ap.add_argument("--load", type=float, default=2.0, help="c4: offered load of the timed streams (x plain capacity)")
    ap.add_argument("--roofline-requests", type=int, default=32, help="c4: requests of the profiled pass")
    ap.add_argument("--no-c3", action="store_true", help="c4: leave out the batched configuration")
    ap.add_argument("--roofline-only", action="store_true",
                    help="c4: only the profiled pass of the roofline (no served streams, no batched run)")
    ap.add_argument("--batch-reps", type=int, default=3, help="c3: timed batches")
    ap.add_argument("--tune-ms", type=int, default=10000, help="c3: autotuner budget (0 = default tuning)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="wall seconds of the CPU baseline sample")
    ap.add_argument("--tune", default="", help="launch tuning overrides, e.g. wg_size=64,blocks_per_cu=16")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="after the run, copy MIB MiB device-to-device three times (hardware-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU (asked for --gpus {args.gpus}): every box here has one GPU, and "
              "sharding over several is a design only.", file=sys.stderr)
        sys.exit(2)
    if args.steps is None:
        args.steps = 20
    import omr

    ctx = omr.Context(0)
    dev = ctx.device_info()
    if args.tune:
        ctx.set_tuning(**{k: int(v) for k, v in (kv.split("=") for kv in args.tune.split(","))})
    if args.workload == "c3":
        r = run_batched(ctx, omr, args)
        out = {"metric": METRIC, "value": r["batch_mrays"], "unit": "Mrays/s", "n_gpus": 1,
               "steps": r["batch_reps"], "warmup": 1, "ms_per_step": r["batch_ms"], "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
               "config": {"workload": r["workload"], "mode": "batched", "device": dev["name"]}}
        out.update({k: v for k, v in r.items() if k != "workload"})
        items = []
    elif args.workload == "c4":
        out, items = run_served(ctx, omr, args, dev)
        if not args.no_c3 and not args.roofline_only:
            out["batched"] = run_batched(ctx, omr, args)
    else:
        out, items = run_plain(ctx, omr, args, dev)
    if not args.no_cpu and items:
        threads = cpu_threads()
        model, logical, usable = host_cpu()
        log(f"cpu baseline: {threads} threads")
        mrays, secs, rows, nreq, rmse = cpu_sample(items, args.cpu_seconds, threads)
        out["cpu_baseline"] = {
            "value": round(mrays, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"bands of 8 rows of the first {nreq} requests' images at their full spp, {rows} rows, "
                      f"{secs:.1f} s wall; binary32 restatement of the reference path with its own binary BVH",
            "cpu_model": model, "host_logical_cpus": logical, "usable_cpus": usable,
            "rmse_vs_gpu": rmse,
        }
        cal = cpu_calibration()
        if cal:
            out["cpu_baseline"]["calibration"] = cal
            out["cpu_baseline"]["reference_equivalent_estimate"] = round(mrays / cal["restatement_over_reference"], 3)
    print(json.dumps(out), flush=True)
    ctx.close()
    if args.calib_copy > 0:
        copy_calibration(args.calib_copy << 20)


def copy_calibration(nbytes):
    """Known-size traffic for scaling hardware byte counters: one fill of
    nbytes and three device-to-device copies of nbytes (each reads and writes
    nbytes once)."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so.7")
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
