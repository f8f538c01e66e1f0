"""Calibrates the CPU restatement (oracle/omr_oracle.c, the CPU baseline that
bench.py times) against the reference's own PathTracer, compiled from its
sources into oracle/_ref/libref.so.  Runs only where the reference checkout
is present; writes oracle/cpu_calibration.json, which bench.py copies into
its cpu_baseline record.

Both render the same scene with the same camera and sample counts in the
reference's mode: Russian roulette after depth 5 and the depth cap of 64
(src/renderer/PathTracer.cpp:177-178,349-360), glass split at depth <= 2.
Workloads:
  c0      the config-0 Cornell box (14 primitives: walls, light quad, mirror
          and glass spheres), 64x64, 2x2 sub-pixels x 64 batches
  shrine  input_data/shrine.obj (970 triangles, textured with shrine.png)
          under input_data/Barce_Rooftop_C_Env.hdr, 64x64, 2x2 x 32 batches
Modes: the reference with its binary BVH (Scene::ConstructBVH, BVH.cpp) and
with brute force over its object list (Scene.cpp:58-70; its QBVH traversal
builds only on Windows, QBVH.cpp:23-27), each on 1 and 2 threads (its row
bands write viewport threadIndex / 2, so more threads split the image,
PathTracer.cpp:95-107); the restatement with its own binary BVH and with the
4-wide BVH the GPU traverses, on 1 and 2 threads.

Rate: segments (closest-hit queries, PathTracer.cpp:183) per wall second.
The reference's own count is private (PathTracer::m_checkIntersectionCount),
so its rate divides the restatement's segment count for the same workload by
the reference's wall time: both run the same estimator, so the expected
number of segments per sample is the same (their random streams differ; the
count's sampling error over ~10^6 segments is well under 1 %).  Images are
compared by per-pixel z-scores against the reference's batch statistics,
over the pixels where the reference's batches vary (a pixel that saw no
light in any of the reference's batches has no variance to scale by; their
share is reported), and by the ratio of the image means.

Usage: python oracle/calibrate_cpu.py [--reps N]
"""
import argparse
import json
import os
import platform
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF_INPUT = "/root/reference/input_data"
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(ROOT, "omochi-renderer3_amd"))

import omr  # noqa: E402
import pyoracle  # noqa: E402
import pyref  # noqa: E402

SHRINE_SCENE = """Base Dir: {base}
IBL: Barce_Rooftop_C_Env.hdr
Space Partitioning: {mode}

Objects:

Obj Mesh {{
	FileName: shrine.obj
	Position: 0, 0, 0
}}
"""


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def shrine_scene(tmp, mode):
    path = os.path.join(tmp, f"shrine_{mode or 'none'}.scene")
    text = SHRINE_SCENE.format(base=REF_INPUT, mode=mode)
    if not mode:
        text = text.replace("Space Partitioning: \n", "")
    with open(path, "w") as f:
        f.write(text)
    return path


def shrine_view(scene):
    """A camera looking at the model from the front, 30 % margin."""
    v = scene.tri[:, 0:3]
    pts = np.concatenate([v, v + scene.tri[:, 3:6], v + scene.tri[:, 6:9]])
    lo, hi = pts.min(0).astype(np.float64), pts.max(0).astype(np.float64)
    c = (lo + hi) / 2
    extent = float(max(hi[0] - lo[0], hi[1] - lo[1]))
    screen_h, dist = 30.0, 40.0
    back = dist * extent * 1.3 / screen_h + (hi[2] - lo[2]) / 2
    pos = (float(c[0]), float(c[1]), float(c[2] + back))
    return dict(pos=pos, dir=(0.0, 0.0, -1.0), screen_h=screen_h, dist=dist)


def zstats(img, mean, var, batches, nb):
    live = var > 0
    se = np.sqrt(var[live] / batches + var[live] / nb)
    z = np.abs(img.astype(np.float64)[live] - mean[live]) / se
    return {"mean_abs_z": round(float(z.mean()), 3), "frac_abs_z_gt_4": round(float((z > 4).mean()), 5),
            "pixels_without_variance": round(float(1.0 - live.mean()), 4),
            "image_mean_ratio": round(float(img.mean() / mean.mean()), 4)}


def run_workload(name, scene, view, w, h, ss, batches, ref_open, reps):
    cam = omr.camera_lookat(w, h, view["pos"], view["dir"], screen_height=view["screen_h"], dist=view["dist"])
    prm = omr.params(ss=ss, spp_end=batches, max_depth=0, rr_depth=5, seed=0x6F6D7233)
    nodes, order, _, _ = omr.build_bvh(scene)
    port = {}
    seg = None
    img = None
    for accel, label in ((pyoracle.OWN_BVH, "binary_bvh"), (pyoracle.WIDE_BVH, "wide4_bvh")):
        o = pyoracle.Oracle(scene, (nodes, order))
        for t in (1, 2):
            best = None
            for _ in range(reps):
                t0 = time.perf_counter()
                im, s, _ = o.render(cam, prm, accel, threads=t)
                dt = time.perf_counter() - t0
                best = dt if best is None else min(best, dt)
            seg = s if seg is None else seg
            assert s == seg, "the restatement's segment count depends only on the seeds"
            img = im
            port[f"{label}_t{t}"] = {"seconds": round(best, 3), "mrays": round(s / best / 1e6, 3)}
        o.close()
    ref = {}
    stats = None
    for mode in ("bvh", "brute"):
        r = ref_open(mode)
        for t in (1, 2):
            best = None
            for _ in range(reps):
                mean, var = r.render(w, h, view["pos"], view["dir"], (0, 1, 0), view["screen_h"], view["dist"], ss,
                                     batches, threads=t)
                dt = pyref.last_seconds()
                best = dt if best is None else min(best, dt)
                if stats is None:
                    stats = zstats(img, mean, var, batches, batches)
            ref[f"{mode}_t{t}"] = {"seconds": round(best, 3), "mrays": round(seg / best / 1e6, 3)}
        r.close()
    ratio = {f"t{t}": round(port[f"binary_bvh_t{t}"]["mrays"] / ref[f"bvh_t{t}"]["mrays"], 3) for t in (1, 2)}
    out = {"workload": name, "image": [w, h], "supersamples": ss, "batches": batches, "segments": int(seg),
           "restatement": port, "reference": ref, "restatement_over_reference_bvh": ratio, "zscores": stats}
    print(json.dumps(out), flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--out", default=os.path.join(HERE, "cpu_calibration.json"))
    args = ap.parse_args()
    if not os.path.isdir(REF_INPUT):
        sys.exit("the reference checkout is not present")
    results = []

    c0 = omr.procedural(0, width=64, height=64)
    view0 = dict(pos=(50.0, 50.0, 260.0), dir=(0.0, 0.0, -1.0), screen_h=30.0, dist=40.0)
    results.append(run_workload("c0", c0, view0, 64, 64, 2, 64,
                                lambda mode: pyref.RefScene(c0, bvh=(mode == "bvh")), args.reps))

    with tempfile.TemporaryDirectory() as tmp:
        here = os.getcwd()
        os.chdir(tmp)
        try:
            shrine = omr.load_scene_file(shrine_scene(tmp, "BVH"))
            view = shrine_view(shrine)
            results.append(run_workload(
                "shrine", shrine, view, 64, 64, 2, 32,
                lambda mode: pyref.RefFileScene(shrine_scene(tmp, "BVH" if mode == "bvh" else "")), args.reps))
        finally:
            os.chdir(here)

    doc = {"cpu_model": cpu_model(), "logical_cpus": os.cpu_count(),
           "restatement_build": "gcc -O2 -std=c11 -ffp-contract=off (oracle/Makefile), binary32",
           "reference_build": "g++ -O2 -std=c++14 (oracle/Makefile), FP64, compiled from its sources",
           "results": results}
    with open(args.out, "w") as f:
        json.dump(doc, f, indent=1)
    print("wrote", args.out)


if __name__ == "__main__":
    main()
