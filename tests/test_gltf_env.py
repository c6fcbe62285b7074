"""LoadScene's environment sources for glTF files (the reference's
src/scene.cpp:307-359, 426-624, 853-859 and pathtracing/scene_data.h:52-69,
with stb_image's stbi_loadf as the decoder): the HDR file named in extras or
extensions (Radiance RGBE, flat and run-length rows, or PNG), the
EXT_lights_image_based cube turned equirectangular, the fallback sky when the
file cannot be read or decoded, and the errors of a bad cube.

The expected images are written here independently: RGBE pixels as
c * 2^(e - 136) in binary32, PNG channels through the C library's powf, and
the cube lookup restated in numpy binary32 with the C library's sinf/cosf, so
every texel must match bit for bit.  Geometric checks (which face the zenith,
the nadir, the horizon and azimuth 0 see) pin the axes independently of the
restatement."""
import ctypes
import ctypes.util
import json

import numpy as np
import pytest

import gltf_files as gf

LIBM = ctypes.CDLL(ctypes.util.find_library("m"))
for _f in ("powf", "sinf", "cosf"):
    getattr(LIBM, _f).restype = ctypes.c_float
LIBM.powf.argtypes = [ctypes.c_float, ctypes.c_float]
LIBM.sinf.argtypes = [ctypes.c_float]
LIBM.cosf.argtypes = [ctypes.c_float]
F = np.float32


from gltf_files import hdr_file


def rgbe_expected(rgbe: np.ndarray) -> np.ndarray:
    e = rgbe[..., 3].astype(np.int64)
    scale = np.where(e > 0, np.ldexp(np.float64(1.0), e - 136), 0.0).astype(F)  # exact powers of two in binary32
    out = np.empty(rgbe.shape, F)
    out[..., :3] = rgbe[..., :3].astype(F) * scale[..., None]
    out[..., 3] = F(1.0)
    return out


def _rgbe_image(rng, h, w):
    img = rng.integers(0, 256, (h, w, 4)).astype(np.uint8)
    img[..., 3] = rng.integers(100, 160, (h, w))
    img[0, 1:6, :] = img[0, 0, :]  # runs
    img[1, :, 3] = 0  # zero exponent: black
    img[2, :, 0] = 7
    img[0, 0, 0] = 7  # first pixel cannot read as an RLE row marker
    return img


# --- scenes ---------------------------------------------------------------------------------------------------------
def _scene(tmp_path, mutate, name="scene.gltf"):
    p = gf.write(tmp_path, "datauri")
    doc = json.loads(p.read_text())
    mutate(doc)
    out = tmp_path / name
    out.write_text(json.dumps(doc))
    return str(out)


def _env(cg, path):
    return cg.LoadScene(path).describe()["env"]


def test_hdr_rle_rows(cg, tmp_path):
    img = _rgbe_image(np.random.default_rng(1), 8, 19)
    (tmp_path / "sky.hdr").write_bytes(hdr_file(img, rle=True))
    env = _env(cg, _scene(tmp_path, lambda d: d.__setitem__("extras", {"environment_hdr": "sky.hdr"})))
    assert env.shape == (8, 19, 4) and env.dtype == np.float32
    assert env.tobytes() == rgbe_expected(img).tobytes()
    assert env[1, :, :3].max() == 0.0 and (env[..., 3] == 1.0).all()
    # a known value: (128, 64, 32, e = 129) is (1, 0.5, 0.25)
    one = np.array([[[128, 64, 32, 129]] * 8], np.uint8)
    (tmp_path / "one.hdr").write_bytes(hdr_file(one, rle=True))
    env = _env(cg, _scene(tmp_path, lambda d: d.__setitem__("extras", {"hdr": "one.hdr"}), "one.gltf"))
    assert np.array_equal(env[0, 0], np.array([1.0, 0.5, 0.25, 1.0], F))


@pytest.mark.parametrize("w", [5, 19])
def test_hdr_flat_rows(cg, tmp_path, w):
    """Narrow images are always flat; a wide one whose first pixel is not a
    run-length marker is read flat from there on (stb's reading)."""
    img = _rgbe_image(np.random.default_rng(2), 6, w)
    (tmp_path / "sky.hdr").write_bytes(hdr_file(img, rle=False, magic=b"#?RGBE"))
    env = _env(cg, _scene(tmp_path, lambda d: d.__setitem__("extras", {"hdr": "sky.hdr"})))
    assert env.tobytes() == rgbe_expected(img).tobytes()


def test_hdr_marker_missing_mid_file(cg, tmp_path):
    """Rows 0-1 run-length encoded, then plain quadruples: the first of them
    becomes pixel 0 of row 0 and the rest fill the image flat from pixel 1."""
    rng = np.random.default_rng(3)
    w, h = 12, 5
    rle_rows = rng.integers(0, 256, (2, w, 4)).astype(np.uint8)
    rle_rows[..., 3] = 130
    flat = rng.integers(0, 256, (w * h, 4)).astype(np.uint8)
    flat[:, 3] = 128
    flat[0, 0] = 9  # not a marker
    data = hdr_file(rle_rows, rle=True, res_line=f"-Y {h} +X {w}\n".encode()) + flat.tobytes()
    (tmp_path / "sky.hdr").write_bytes(data)
    env = _env(cg, _scene(tmp_path, lambda d: d.__setitem__("extras", {"hdr": "sky.hdr"})))
    exp = rgbe_expected(flat.reshape(h, w, 4))
    assert env.tobytes() == exp.tobytes()


def test_hdr_header_forms(cg, tmp_path):
    img = _rgbe_image(np.random.default_rng(4), 3, 9)
    long_line = b"# " + b"x" * 3000 + b"\n"
    (tmp_path / "sky.hdr").write_bytes(hdr_file(img, rle=True, extra_header=long_line,
                                                res_line=b"-Y 3   +X 9\n"))
    env = _env(cg, _scene(tmp_path, lambda d: d.__setitem__("extras", {"hdr": "sky.hdr"})))
    assert env.tobytes() == rgbe_expected(img).tobytes()


@pytest.mark.parametrize("variant", ["no_format", "bad_layout", "truncated", "bad_run", "missing", "jpeg", "empty"])
def test_undecodable_hdr_keeps_the_fallback_sky(cg, tmp_path, variant):
    img = _rgbe_image(np.random.default_rng(5), 4, 16)
    good = hdr_file(img, rle=True)
    data = {"no_format": good.replace(b"FORMAT=32-bit_rle_rgbe", b"FORMAT=32-bit_rle_xyze"),
            "bad_layout": good.replace(b"-Y 4 +X 16", b"+Y 4 +X 16"),
            "truncated": good[:-5],
            "bad_run": good.replace(bytes([2, 2, 0, 16]), bytes([2, 2, 0, 16, 200]), 1),
            "jpeg": b"\xff\xd8\xff\xe0" + b"\0" * 64,
            "empty": b""}.get(variant)
    if data is not None:
        (tmp_path / "sky.hdr").write_bytes(data)
    s = cg.LoadScene(_scene(tmp_path, lambda d: d.__setitem__("extras", {"hdr": "sky.hdr"})))
    assert s.describe()["env"] is None
    assert s.scene_data().arrays()["env"].shape == (256, 512, 4)  # the fallback sky


def test_png_environment(cg, tmp_path):
    rng = np.random.default_rng(6)
    px = rng.integers(0, 256, (3, 5, 4))
    px[0, 0] = [0, 255, 128, 77]
    (tmp_path / "sky.png").write_bytes(gf.png(px, 6))
    env = _env(cg, _scene(tmp_path, lambda d: d.__setitem__("extras", {"hdr": "sky.png"})))
    exp = np.empty(px.shape, F)
    for idx in np.ndindex(px.shape[:2]):
        for c in range(3):
            exp[idx + (c,)] = LIBM.powf(F(px[idx + (c,)]) / F(255.0), F(2.2))
        exp[idx + (3,)] = F(px[idx + (3,)]) / F(255.0)
    assert env.tobytes() == exp.tobytes()
    assert env[0, 0, 0] == 0.0 and env[0, 0, 1] == 1.0 and abs(env[0, 0, 2] - (128 / 255) ** 2.2) < 1e-6


def _solid_hdr(tmp_path, name, value):
    img = np.zeros((1, 8, 4), np.uint8)
    img[..., 0] = value
    img[..., 3] = 136  # scale 1
    (tmp_path / name).write_bytes(hdr_file(img, rle=True))


@pytest.mark.parametrize("case,expect", [
    ({"extras": {"hdr": "a.hdr"}, "scenes": [{"nodes": gf.SCENE_NODES, "extras": {"environment_hdr": "b.hdr"}}]}, 10),
    ({"extras": {"hdr": "a.hdr", "environmentMap": "c.hdr"}}, 30),
    ({"extras": {"environment_hdr": "", "hdr": "a.hdr"}}, 10),
    ({"extras": {"settings": {"lighting": [{"hdrFile": "my%20sky.hdr"}]}}}, 40),
    ({"extensions": {"KHR_lights_punctual": {"hdr": "a.hdr"}, "EXT_custom": {"environmentHdr": "c.hdr"}}}, 30),
    ({"scenes": [{"nodes": gf.SCENE_NODES, "extensions": {"EXT_custom": {"environmentHDR": "b.hdr"}}}]}, 20),
    ({"extras": {"hdr": 5}}, None),
])
def test_hdr_path_sources_and_keys(cg, tmp_path, case, expect):
    for name, v in (("a.hdr", 10), ("b.hdr", 20), ("c.hdr", 30), ("my sky.hdr", 40)):
        _solid_hdr(tmp_path, name, v)
    env = _env(cg, _scene(tmp_path, lambda d: d.update(case)))
    if expect is None:
        assert env is None
    else:
        assert env[0, 0, 0] == expect


# --- EXT_lights_image_based ---------------------------------------------------------------------------------------
def _faces(n=4):
    """Six n x n faces: red = face index, green = x, blue = y, as RGBE (exact)."""
    out = []
    for f in range(6):
        img = np.zeros((n, n, 4), np.uint8)
        img[..., 0] = 16 * (f + 1)
        img[..., 1] = np.arange(n)[None, :] * 8 + 1
        img[..., 2] = np.arange(n)[:, None] * 8 + 1
        img[..., 3] = 136
        out.append(img)
    return out


def _rotate(q, d):
    im = np.array(q[:3], F)
    re = F(q[3])
    norm = ((im[0] * im[0] + im[1] * im[1]) + im[2] * im[2]) + re * re
    if norm <= 0:
        return d
    sn = np.sqrt(norm)
    im = im / sn
    re = re / sn
    cross = lambda a, b: np.array([a[1] * b[2] - b[1] * a[2], a[2] * b[0] - b[2] * a[0], a[0] * b[1] - b[0] * a[1]], F)
    t = F(2.0) * cross(im, d)
    return (d + re * t) + cross(im, t)


def cube_expected(faces: list[np.ndarray], rot) -> np.ndarray:
    """The reference's CubeToEnvironmentTexture (src/scene.cpp:541-583) in binary32."""
    fl = [rgbe_expected(f) for f in faces]
    n = faces[0].shape[0]
    W, H = 4 * n, 2 * n
    inv = [-rot[0], -rot[1], -rot[2], rot[3]]
    pi = F(np.pi)
    out = np.zeros((H, W, 4), F)
    for y in range(H):
        v = (F(y) + F(0.5)) / F(H)
        theta = v * pi
        st, z = F(LIBM.sinf(theta)), F(LIBM.cosf(theta))
        for x in range(W):
            u = (F(x) + F(0.5)) / F(W)
            phi = ((u - F(0.5)) * F(2.0)) * pi
            e = np.array([F(LIBM.cosf(phi)) * st, F(LIBM.sinf(phi)) * st, z], F)
            d = _rotate(inv, np.array([e[0], -e[2], e[1]], F))
            ax, ay, az = abs(d[0]), abs(d[1]), abs(d[2])
            if ax >= ay and ax >= az:
                face = 1 if d[0] >= 0 else 0
                cu = (-d[2] / ax + F(1)) * F(0.5) if d[0] >= 0 else (d[2] / ax + F(1)) * F(0.5)
                cv = (-d[1] / ax + F(1)) * F(0.5)
            elif ay >= az:
                face = 2 if d[1] >= 0 else 3
                cu = (d[0] / ay + F(1)) * F(0.5)
                cv = (d[2] / ay + F(1)) * F(0.5) if d[1] >= 0 else (-d[2] / ay + F(1)) * F(0.5)
            else:
                face = 4 if d[2] >= 0 else 5
                cu = (d[0] / az + F(1)) * F(0.5) if d[2] >= 0 else (-d[0] / az + F(1)) * F(0.5)
                cv = (F(1) - d[1] / az) * F(0.5)
            tu = min(max(F(1) - cu, F(0)), F(1))
            tv = min(max(F(1) - cv, F(0)), F(1))
            tx = min(n - 1, int(tu * F(n)))
            ty = min(n - 1, int(tv * F(n)))
            out[y, x, :3] = fl[face][ty, tx, :3]
            out[y, x, 3] = 1.0
    return out


def _ibl_scene(tmp_path, faces_bytes, ibl, name="ibl.gltf", extra=None):
    def mutate(d):
        base = len(d["images"])
        import base64
        for b in faces_bytes:
            d["images"].append({"uri": "data:image/vnd.radiance;base64," + base64.b64encode(b).decode()})
        d.setdefault("extensions", {})["EXT_lights_image_based"] = ibl(base)
        d["extensionsUsed"] = ["EXT_lights_image_based"]
        if extra:
            d.update(extra)
    return _scene(tmp_path, mutate, name)


@pytest.mark.parametrize("rot", [[0.0, 0.0, 0.0, 1.0], [0.0, 0.7071068, 0.0, 0.7071068], [0.2, -0.4, 0.1, 0.9]])
def test_ibl_cube_to_equirect(cg, tmp_path, rot):
    faces = _faces()
    data = [hdr_file(f, rle=False) for f in faces]
    lights = lambda b: {"lights": [{"name": "sky", "rotation": rot, "intensity": 1.0,
                                    "specularImages": [[b + i for i in range(6)], [0, 0, 0, 0, 0, 0]]}]}
    env = _env(cg, _ibl_scene(tmp_path, data, lights))
    exp = cube_expected(faces, [F(r) for r in rot])
    assert env.shape == exp.shape == (8, 16, 4)
    assert env.tobytes() == exp.tobytes()
    if rot == [0.0, 0.0, 0.0, 1.0]:
        red = (env[..., 0] / 16).round().astype(int) - 1  # face index of each texel
        # The reference takes engine (x, y, z) to glTF (x, -z, y) (src/scene.cpp:522-524), so the zenith
        # row sees the -y face (3) and the nadir row the +y face (2).
        assert (red[0] == 3).all() and (red[-1] == 2).all()
        ring = set(red[3].tolist()) | set(red[4].tolist())
        assert ring == {0, 1, 4, 5}  # the horizon sees the four side faces
        # azimuth 0 (engine +x, the texel column just right of the middle) is glTF +x: face 1
        assert red[4, 8] == 1 and red[4, 0] == 0


def test_ibl_wins_over_hdr_and_png_faces(cg, tmp_path):
    _solid_hdr(tmp_path, "a.hdr", 10)
    rng = np.random.default_rng(8)
    pngs = [rng.integers(0, 256, (2, 2, 4)) for _ in range(6)]
    lights = lambda b: {"lights": [{"specularImages": [[b, b + 1, b + 2, b + 3, b + 4, b + 5]]}]}
    env = _env(cg, _ibl_scene(tmp_path, [gf.png(p, 6) for p in pngs], lights, extra={"extras": {"hdr": "a.hdr"}}))
    assert env.shape == (4, 8, 4) and (env[..., 3] == 1.0).all()
    vals = {float(LIBM.powf(F(v) / F(255.0), F(2.2))) for p in pngs for v in p[..., 0].ravel()}
    assert set(env[..., 0].ravel().tolist()) <= vals


@pytest.mark.parametrize("sets", [[[0, 1, 2, 3, 4]], [[0, 1, 2, 3, 4, -5]], [[0, 1, 2, 3, 4, 5.5]], [], 3])
def test_malformed_specular_images_mean_no_cube(cg, tmp_path, sets):
    _solid_hdr(tmp_path, "a.hdr", 10)
    env = _env(cg, _ibl_scene(tmp_path, [], lambda b: {"lights": [{"specularImages": sets}]},
                              extra={"extras": {"hdr": "a.hdr"}}))
    assert env[0, 0, 0] == 10  # no cube: the HDR file


@pytest.mark.parametrize("kind,msg", [("index", "invalid image"), ("size", "share one size"), ("jpeg", "decode")])
def test_ibl_errors(cg, tmp_path, kind, msg):
    faces = _faces()
    data = [hdr_file(f, rle=False) for f in faces]
    if kind == "size":
        data[3] = hdr_file(_faces(2)[3], rle=False)
    if kind == "jpeg":
        data[5] = b"\xff\xd8\xff\xe0" + b"\0" * 64
    off = 100 if kind == "index" else 0
    lights = lambda b: {"lights": [{"specularImages": [[b + off + i for i in range(6)]]}]}
    with pytest.raises(cg.NativeError, match=msg):
        cg.LoadScene(_ibl_scene(tmp_path, data, lights))


def test_tag_follows_the_environment(cg, tmp_path):
    _solid_hdr(tmp_path, "a.hdr", 10)
    p = _scene(tmp_path, lambda d: d.__setitem__("extras", {"hdr": "a.hdr"}))
    t1 = cg.LoadScene(p).describe()["tag"]
    _solid_hdr(tmp_path, "a.hdr", 11)
    t2 = cg.LoadScene(p).describe()["tag"]
    assert t1 != t2


def test_environment_scene_prep_matches_oracle(cg, oracle, tmp_path):
    img = _rgbe_image(np.random.default_rng(9), 16, 32)
    (tmp_path / "sky.hdr").write_bytes(hdr_file(img, rle=True))
    s = cg.LoadScene(_scene(tmp_path, lambda d: d.__setitem__("extras", {"hdr": "sky.hdr"})))
    a = s.scene_data(share_blas=False).arrays()
    b = oracle.OracleScene.build(s.describe()).arrays()
    for k in a:
        assert a[k].tobytes() == b[k].tobytes(), k
    assert a["env"].tobytes() == rgbe_expected(img).tobytes()
