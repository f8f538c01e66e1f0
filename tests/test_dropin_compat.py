"""The C++ drop-in accepts the reference's own scene code with only its
#include lines changed, and its host-side FP64 intersection code gives the
reference's hit records.

oracle/Makefile's dropin_compat target compiles the reference's
CornellBoxScene.h / .cpp (src/scenes/), #include lines removed, against
OmochiRenderer.h and libomr.so, with the caller in tests/dropin/compat_main.cpp
(the render path of the reference's main.cpp plus probes of the interfaces a
caller uses).  It builds only where the reference checkout is present; the
program itself travels with the built libraries.  Without a GPU,
Scene::CheckIntersection tests every object in FP64 on the host, which the
test compares with the reference's own hits; tests/test_gpu_dropin.py runs
the same program on the GPU."""
import os
import subprocess

import numpy as np
import pytest

import golden_inputs as gi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src/scenes"
LIB = os.path.join(ROOT, "omochi-renderer3_amd", "lib")
GOLDEN = os.path.join(ROOT, "tests", "golden")
EXE = os.path.join(ROOT, "oracle", "_ref", "dropin_compat")


@pytest.fixture(scope="module")
def program(tmp_path_factory):
    if not os.path.exists(os.path.join(REF, "CornellBoxScene.cpp")):
        pytest.skip("reference checkout not present")
    if not os.path.exists(os.path.join(LIB, "libomr.so")):
        pytest.skip("library not built")
    p = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "_ref/dropin_compat"], capture_output=True,
                       text=True)
    assert p.returncode == 0, (p.stdout + p.stderr)[-4000:]
    return EXE, tmp_path_factory.mktemp("compat")


def test_reference_scene_code_compiles_against_dropin(program):
    """CornellBoxScene.h / .cpp as the reference has them, #include lines
    aside, plus main.cpp's render path, build against OmochiRenderer.h and
    link against libomr.so."""
    exe, _ = program
    assert os.path.exists(exe)


def test_dropin_host_hits_equal_reference(program):
    """The drop-in's Scene::CheckIntersection on the reference's own scene
    code (host FP64 when no GPU is present) against the reference's hits on
    the same rays: same distances (1e-12 relative), normals, materials and
    rates; and the small interfaces give the reference's values."""
    exe, d = program
    import omr

    here = os.getcwd()
    os.chdir(os.path.join(GOLDEN, "scenes"))
    try:
        scene = omr.load_scene_file("cornell_ref.scene")
    finally:
        os.chdir(here)
    rays = gi.cornell_rays(scene)
    rays.astype(np.float32).tofile(d / "rays.bin")
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1")  # host path even where a GPU is present
    p = subprocess.run([exe, str(d / "rays.bin"), str(d / "hits.bin")], capture_output=True, text=True, env=env,
                       timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    lines = {ln.split()[0]: ln.split()[1:] for ln in p.stdout.splitlines()}
    assert lines["lights"][0] == "1"
    assert lines["lights"][1] == "ray" and lines["lights"][5] == "const"
    ray_dir = np.array([float(x) for x in lines["lights"][2:5]])
    assert abs(np.linalg.norm(ray_dir) - 1.0) < 1e-12
    c = [float(x) for x in lines["lights"][6:]]
    assert c[0] == (float(np.finfo(np.float32).max) - 1) / 2 and c[1] == np.pi
    assert c[2:] == [0.0005, 1.0, 1.5]
    assert lines["viewport"] == ["7", "8", "3"]
    rec = np.fromfile(d / "hits.bin", np.float64).reshape(-1, 9)
    g = np.load(os.path.join(GOLDEN, "hits_cornell_ref.npz"))
    hit = g["obj"] >= 0
    assert np.array_equal(rec[:, 0] >= 0, hit)
    assert np.allclose(rec[hit, 0], g["t"][hit], rtol=1e-12, atol=0)
    assert np.allclose(rec[hit, 1:4], g["normal"][hit], rtol=0, atol=1e-6)
    assert np.array_equal(rec[hit, 4], g["nmat"][hit].astype(np.float64))
    m0 = g["mat"][hit][:, 0]
    assert np.array_equal(rec[hit, 5], m0[:, 0].astype(np.float64))
    assert np.allclose(rec[hit, 6], m0[:, 1], rtol=1e-6)
    assert np.allclose(rec[hit, 7], m0[:, 4], rtol=1e-6)
    assert np.allclose(rec[hit, 8], m0[:, 8], rtol=1e-6)
