"""Scene-side readers against the reference: the RGBE (.hdr) reader and
writer, the .scene / OBJ / MTL loaders, and the infinite-floor primitive.

Fixtures in tests/golden/ were written by tests/golden/make_golden.py from
the reference's own HDRImage, SceneFromExternalFile, Model and InfiniteFloor
compiled in place; rgbe_*.hdr and scenes/sky.hdr are files the reference's
writer produced.  The loaders are host code (no GPU); the closest hits of a
loaded scene go through the CPU restatement, which the GPU path matches bit
for bit elsewhere.
"""
import hashlib
import os

import numpy as np
import pytest

import golden_inputs as gi

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SCENES = os.path.join(GOLDEN, "scenes")


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _rgbe_pixels(path):
    """Pixel bytes of a flat (not run-length encoded) Radiance file."""
    raw = open(path, "rb").read()
    head_end = raw.index(b"\n\n") + 2
    size_end = raw.index(b"\n", head_end) + 1
    h, w = (int(x) for x in raw[head_end:size_end].split()[1::2])
    return w, h, np.frombuffer(raw[size_end:], np.uint8).reshape(h, w, 4)


def _rgbe_decode(px):
    """Radiance RGBE: a shared exponent byte e and 8-bit mantissas m give
    m * 2^(e - 136); e = 0 is black."""
    e = px[..., 3].astype(np.int32)
    f = np.where(e > 0, np.ldexp(1.0, e - 136), 0.0)
    return (px[..., :3].astype(np.float64) * f[..., None]).astype(np.float32)


def test_rgbe_known_pixels(tmp_path):
    """Closed forms of the format: (128, 64, 32, e=129) is (1, 1/2, 1/4);
    e = 0 is black whatever the mantissas; writing (1, 0.5, 0.25) and 255*2^k
    gives the same bytes back."""
    import omr

    p = tmp_path / "k.hdr"
    body = bytes([128, 64, 32, 129, 7, 9, 11, 0, 255, 0, 1, 136])
    p.write_bytes(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 1 +X 3\n" + body)
    img = omr.hdr_read(str(p))
    assert img.shape == (1, 3, 3)
    assert img[0].tolist() == [[1.0, 0.5, 0.25], [0.0, 0.0, 0.0], [255.0, 0.0, 1.0]]
    omr.hdr_write(str(tmp_path / "o.hdr"), img)
    w, h, px = _rgbe_pixels(str(tmp_path / "o.hdr"))
    assert (w, h) == (3, 1)
    assert px.tobytes() == bytes([128, 64, 32, 129, 0, 0, 0, 0, 255, 0, 1, 136])


@pytest.mark.parametrize("name", sorted(gi.HDR_SETS))
def test_rgbe_writer_bytes_match_reference(name, tmp_path):
    """HDRImage::WriteToRadianceFile (src/tools/HDRImage.cpp:188-290): same
    header, same RGBE rounding, same run-length scanlines, byte for byte."""
    import omr

    img = gi.hdr_image(*gi.HDR_SETS[name])
    g = np.load(os.path.join(GOLDEN, "rgbe_decoded.npz"))
    assert str(g[f"{name}_input_sha256"]) == _sha(img), "input image drifted from the fixture"
    out = tmp_path / "w.hdr"
    omr.hdr_write(str(out), img)
    assert out.read_bytes() == open(os.path.join(GOLDEN, f"rgbe_{name}.hdr"), "rb").read()


def test_sky_writer_bytes_match_reference(tmp_path):
    import omr

    sky = gi.sky_image()
    out = tmp_path / "sky.hdr"
    omr.hdr_write(str(out), sky)
    assert out.read_bytes() == open(os.path.join(SCENES, "sky.hdr"), "rb").read()


@pytest.mark.parametrize("name", ["rle", "sky"])
def test_rgbe_reader_matches_reference(name):
    """HDRImage::ReadFromRadianceFile (src/tools/HDRImage.cpp:46-186) on
    run-length encoded files: the same floats, bit for bit."""
    import omr

    g = np.load(os.path.join(GOLDEN, "rgbe_decoded.npz"))
    path = os.path.join(SCENES, "sky.hdr") if name == "sky" else os.path.join(GOLDEN, f"rgbe_{name}.hdr")
    img = omr.hdr_read(path)
    ref = g[f"{name}_decoded"]
    assert img.shape == ref.shape
    assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))


def test_rgbe_reader_narrow_flat_file():
    """Images narrower than 8 pixels are stored flat.  The reference reader
    passes such a file to its flat reader with the width and height in place
    of the pixel count and start offset (src/tools/HDRImage.cpp:112-115,
    :96-106), so its image holds `width` pixels instead of width x height;
    this reader decodes every pixel: the format's closed form on the bytes
    the reference writer produced."""
    import omr

    path = os.path.join(GOLDEN, "rgbe_flat.hdr")
    w, h, px = _rgbe_pixels(path)
    img = omr.hdr_read(path)
    assert img.shape == (h, w, 3)
    assert np.array_equal(img, _rgbe_decode(px))
    src = gi.hdr_image(*gi.HDR_SETS["flat"])
    big = src.max(axis=2, keepdims=True)
    # 8-bit mantissas of a shared exponent: each channel within 2^-7 of the largest (truncation)
    assert np.all(np.abs(img - src) <= big * 2.0 ** -7 + 1e-30)
    assert img.any()


def test_rgbe_round_trip_rle(tmp_path):
    """Write then read: scanlines of every run shape come back as the
    format's quantization of the input (run-length path, widths 8..300)."""
    import omr

    rng = np.random.default_rng(3)
    for w in (8, 9, 31, 127, 128, 129, 300):
        img = gi.hdr_image(w, 3, int(rng.integers(1 << 30)))
        p = tmp_path / f"r{w}.hdr"
        omr.hdr_write(str(p), img)
        back = omr.hdr_read(str(p))
        big = img.max(axis=2, keepdims=True)
        assert np.all(np.abs(back - img) <= big * 2.0 ** -7 + 1e-30), w
        assert np.all(back <= img * (1 + 2.0 ** -23)), "RGBE truncates"


def test_rgbe_reader_rejects_garbage(tmp_path):
    import omr

    p = tmp_path / "bad.hdr"
    p.write_bytes(b"P6\n3 2\n255\n" + bytes(18))
    with pytest.raises(omr.OmrError):
        omr.hdr_read(str(p))
    with pytest.raises(omr.OmrError):
        omr.hdr_read(str(tmp_path / "missing.hdr"))


# -------------------------------------------------------------- planes

def test_plane_hits_match_reference():
    """InfiniteFloor (src/renderer/InfiniteFloor.h:18-40) tested after the
    other objects: identical primitive ids (planes numbered after spheres)
    and t within 1e-5 relative of the FP64 reference."""
    import pyoracle

    scene = gi.plane_scene()
    kind, n, seed = gi.PLANE_HITS
    rays = gi.hit_rays(scene, n, seed)
    g = np.load(os.path.join(GOLDEN, "hits_planes.npz"))
    assert str(g["rays_sha256"]) == _sha(rays)
    prim, t = pyoracle.Oracle(scene).trace(rays, pyoracle.OWN_BVH)
    assert np.array_equal(prim, g["prim"])
    first = scene.tri.shape[0] + scene.sph.shape[0]
    assert (prim >= first).sum() > 1000 and (prim == first + 1).sum() > 100
    hit = prim >= 0
    rel = np.abs(t[hit].astype(np.float64) - g["t"][hit]) / g["t"][hit]
    assert rel.max() < 1e-5


def test_plane_closed_form():
    """t = (y - o.y) / d.y for |d.y| >= EPS and t >= EPS, else a miss."""
    import ctypes as C

    import pyoracle

    lib = pyoracle.load()
    cases = [((0, 10, 0), (0, -1, 0), 2.5, 7.5), ((1, -3, 2), (0.6, 0.8, 0), 5.0, 10.0),
             ((0, 1, 0), (1, 0, 0), 0.0, None), ((0, 1, 0), (0, 1, 0), 0.0, None),
             ((0, 1, 0), (0, 0.0004, 1), 2.0, None), ((0, 1.0002, 0), (0, -1, 0), 1.0, None)]
    for o, d, y, want in cases:
        r = np.array([*o, 0, *d, 0], np.float32)
        t = C.c_float()
        hit = lib.orc_intersect_plane(r.ctypes.data_as(C.c_void_p), y, C.byref(t))
        if want is None:
            assert not hit, (o, d, y)
        else:
            assert hit and abs(t.value - want) <= 1e-6 * want, (o, d, y, t.value)


# -------------------------------------------------------------- scene files

@pytest.fixture(scope="module")
def loaded():
    import omr

    cwd = os.getcwd()
    os.chdir(SCENES)  # Base Dir and file names resolve from the working directory
    try:
        return omr.load_scene_file("loader.scene")
    finally:
        os.chdir(cwd)


def _normals(scene, prim, rays, t):
    """Shading normal of each hit from the loaded arrays: interpolated
    vertex normals on triangles, radial on spheres, +-y on planes."""
    nt, ns = scene.tri.shape[0], scene.sph.shape[0]
    o = rays[:, 0:3].astype(np.float64)
    d = rays[:, 4:7].astype(np.float64)
    p = o + d * t[:, None]
    out = np.zeros((len(prim), 3))
    tri = prim < nt
    if tri.any():
        T = scene.tri[prim[tri]].astype(np.float64)
        e1, e2, q = T[:, 3:6], T[:, 6:9], p[tri] - T[:, 0:3]
        a11, a12, a22 = (e1 * e1).sum(1), (e1 * e2).sum(1), (e2 * e2).sum(1)
        b1, b2 = (q * e1).sum(1), (q * e2).sum(1)
        det = a11 * a22 - a12 * a12
        u, v = (b1 * a22 - b2 * a12) / det, (b2 * a11 - b1 * a12) / det
        N = scene.tri_n[prim[tri]].astype(np.float64)
        out[tri] = N[:, 0:3] + u[:, None] * N[:, 3:6] + v[:, None] * N[:, 6:9]
    sph = (prim >= nt) & (prim < nt + ns)
    if sph.any():
        out[sph] = p[sph] - scene.sph[prim[sph] - nt, 0:3]
    pl = prim >= nt + ns
    out[pl, 1] = np.where(d[pl, 1] < 0, 1.0, -1.0)
    return out / np.linalg.norm(out, axis=1, keepdims=True)


def test_scene_file_matches_reference(loaded):
    """SceneFromExternalFile (src/scenes/SceneFromExternalFile.cpp) with
    Model::ReadFromObj / LoadMaterialFile (src/renderer/Model.cpp) on a file
    that uses every block type: same hits, distances, normals, materials
    (type, colour, emission, index, normalized rates) and the same object
    order within each primitive kind."""
    import pyoracle

    scene = loaded
    g = np.load(os.path.join(GOLDEN, "hits_loader.npz"))
    rays = gi.loader_rays()
    assert str(g["rays_sha256"]) == _sha(rays)
    nt, ns, npl = scene.tri.shape[0], scene.sph.shape[0], scene.planes.shape[0]
    assert nt + ns + npl == int(g["n_objects"])
    prim, t = pyoracle.Oracle(scene).trace(rays, pyoracle.OWN_BVH)
    obj = g["obj"]
    assert np.array_equal(prim >= 0, obj >= 0)
    hit = prim >= 0
    assert hit.mean() > 0.5

    # the reference lists objects in file order; primitives here are grouped
    # triangles, spheres, planes with file order kept inside each group
    kind = np.where(prim < nt, 0, np.where(prim < nt + ns, 1, 2))
    pairs = sorted(set(zip(prim[hit].tolist(), obj[hit].tolist())))
    assert len({p for p, _ in pairs}) == len(pairs) == len({q for _, q in pairs}), "not one-to-one"
    for k in range(3):
        sel = [(p, q) for p, q in pairs if kind[prim == p][0] == k]
        assert [q for _, q in sel] == sorted(q for _, q in sel)
    assert len(pairs) == int(g["n_objects"]), "every object is hit by some ray"

    rel = np.abs(t[hit].astype(np.float64) - g["t"][hit]) / g["t"][hit]
    assert rel.max() < 1e-5
    n = _normals(scene, prim[hit], rays[hit], g["t"][hit])
    assert np.abs(n - g["normal"][hit]).max() < 1e-4

    pm = scene.prim_mat[prim[hit]]
    cnt = np.ones_like(pm) if scene.prim_mat_count is None else scene.prim_mat_count[prim[hit]]
    assert np.array_equal(cnt, g["nmat"][hit])
    ref = g["mat"][hit]
    for k in range(int(cnt.max())):
        has = cnt > k
        m = scene.mats[pm[has] + k]
        r = ref[has, k]
        assert np.array_equal(m["type"].astype(np.float32), r[:, 0])
        assert np.array_equal(m["color"].astype(np.float32), r[:, 1:4])
        assert np.array_equal(m["emission"].astype(np.float32), r[:, 4:7])
        glass = m["type"] == 2
        assert np.array_equal(m["ior"][glass].astype(np.float32), r[glass, 7])
        assert np.allclose(m["rate"], r[:, 8], rtol=1e-6, atol=0)


def test_scene_file_contents(loaded):
    """What the test file declares: 9 mesh triangles (fans of a quad and a
    pentagon), 6 floor triangles, 3 spheres, one infinite floor at y = -5,
    the 64 x 32 environment and the normalized material rates."""
    scene = loaded
    assert scene.tri.shape[0] == 15 and scene.sph.shape[0] == 3
    assert scene.planes.tolist() == [-5.0]
    assert scene.env is not None and scene.env.shape == (32, 64, 3)
    sky = np.load(os.path.join(GOLDEN, "rgbe_decoded.npz"))["sky_decoded"]
    assert np.array_equal(scene.env, sky)
    s0 = scene.prim_mat[scene.tri.shape[0]]
    assert scene.prim_mat_count[scene.tri.shape[0]] == 2
    assert np.allclose(scene.mats["rate"][s0:s0 + 2], [0.6, 0.4])
    light = scene.mats[scene.prim_mat[scene.tri.shape[0] + 2]]
    assert light["emission"].tolist() == [20.0, 18.0, 15.0]


def test_obj_loader_transform():
    """omr_obj_load places a mesh as Polygon::Transform does
    (src/renderer/Polygon.h:51-61): rotation Rz*Ry*Rx, then the per-axis
    scale along world axes, then the translation; normals rotated only."""
    import omr

    plain = omr.load_obj(os.path.join(SCENES, "mesh.obj"))
    pos, sc, rot = (10.0, 5.0, -20.0), (2.0, 3.0, 1.5), (0.3, -0.7, 1.1)
    moved = omr.load_obj(os.path.join(SCENES, "mesh.obj"), position=pos, scale=sc, rotation=rot)
    assert plain.tri.shape == moved.tri.shape == (9, 9)

    def R(a, b, c):
        cx, sx, cy, sy, cz, sz = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(c), np.sin(c)
        rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
        ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
        rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
        return rz @ ry @ rx

    M = R(*rot)

    def verts(s):
        v0 = s.tri[:, 0:3].astype(np.float64)
        return np.stack([v0, v0 + s.tri[:, 3:6], v0 + s.tri[:, 6:9]], 1)

    want = (verts(plain) @ M.T) * np.array(sc) + np.array(pos)
    assert np.abs(verts(moved) - want).max() < 1e-4
    for k in range(3):
        n = plain.tri_n[:, 3 * k:3 * k + 3].astype(np.float64) @ M.T
        assert np.abs(moved.tri_n[:, 3 * k:3 * k + 3] - n).max() < 1e-5


def test_loader_errors():
    import omr

    with pytest.raises(omr.OmrError):
        omr.load_scene_file(os.path.join(SCENES, "does_not_exist.scene"))
    with pytest.raises(omr.OmrError):
        omr.load_obj(os.path.join(SCENES, "does_not_exist.obj"))
