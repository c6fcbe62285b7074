"""JPEG decoding for glTF textures (crystalgraphics_amd/csrc/host/jpeg.cpp),
pinned by files from an independent encoder (tests/jpeg_files.py, written from
ITU-T T.81): the decoded samples of exactly known quantised coefficients equal
the FP64 inverse DCT (T.81 A.3.3) rounded; a progressive file (spectral
selection and successive approximation) decodes to the same pixels as the
sequential file of the same coefficients; restart intervals change nothing;
near-lossless files reproduce their images; subsampled chroma is upsampled as
declared; malformed and unsupported files are errors; a glTF scene with JPEG
textures loads them.  The reference decodes with stb_image (fetched at build
time, absent here), so stb's own rounding is not pinned: upsampling, colour
conversion and IDCT rounding may differ from it by a few levels."""
import base64
import ctypes as C
import json

import numpy as np
import pytest

import gltf_files as gf
import jpeg_files as jf


def decode(cg, data: bytes) -> np.ndarray:
    from crystalgraphics_amd import native
    h = native.host()
    f = h.cgh_decode_jpeg
    f.restype = C.c_int
    f.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.c_void_p, C.c_size_t,
                  C.POINTER(C.c_char_p)]
    w, hh, err = C.c_uint32(), C.c_uint32(), C.c_char_p()
    if f(data, len(data), C.byref(w), C.byref(hh), None, 0, C.byref(err)) != 0:
        raise ValueError(err.value.decode())
    out = np.zeros((hh.value, w.value, 4), np.uint8)
    assert f(data, len(data), C.byref(w), C.byref(hh), out.ctypes.data, out.nbytes, C.byref(err)) == 0
    return out


def _grey_from_coefficients(comps, W, H):
    """FP64 inverse DCT of the blocks, level shift, round half up, clamp."""
    c = comps[0]
    img = np.zeros((c["blocks"].shape[0] * 8, c["blocks"].shape[1] * 8))
    for by in range(c["blocks"].shape[0]):
        for bx in range(c["blocks"].shape[1]):
            img[by * 8:by * 8 + 8, bx * 8:bx * 8 + 8] = jf.idct(c["blocks"][by, bx].astype(np.float64))
    return np.clip(np.floor(img + 128.0 + 0.5), 0, 255)[:H, :W].astype(np.uint8)


@pytest.mark.parametrize("progressive", [False, True])
def test_samples_are_the_exact_inverse_dct(cg, progressive):
    """Grey image from chosen coefficients: DC only, single basis functions,
    a few random ones, saturating ones; no quantisation error to hide behind."""
    rng = np.random.default_rng(1)
    W, H = 40, 24
    blocks = np.zeros((3, 5, 8, 8), np.int64)
    blocks[0, 0, 0, 0] = 80                      # flat
    blocks[0, 1, 0, 1] = 40                      # one horizontal cosine
    blocks[0, 2, 3, 0] = -35                     # one vertical cosine
    blocks[0, 3, 7, 7] = 90                      # the highest frequency
    blocks[0, 4, 0, 0] = 2000                    # saturates high
    blocks[1, 0, 0, 0] = -2000                   # saturates low
    for by, bx in [(1, 1), (1, 2), (1, 3), (1, 4), (2, 0), (2, 1), (2, 2), (2, 3), (2, 4)]:
        m = rng.random((8, 8)) < 0.3
        blocks[by, bx] = np.where(m, rng.integers(-60, 61, (8, 8)), 0)
    comps = [{"h": 1, "v": 1, "tq": 0, "blocks": blocks, "bw": 5, "bh": 3}]
    data = jf.encode(W, H, comps, [np.ones((8, 8))], progressive=progressive)
    got = decode(cg, data)
    expect = _grey_from_coefficients(comps, W, H)
    assert got.shape == (H, W, 4) and np.all(got[..., 3] == 255)
    for ch in range(3):
        assert np.array_equal(got[..., ch], expect)


def _photo(W, H, seed=3):
    """A smooth colour image with an edge and some texture."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:H, 0:W]
    r = 128 + 100 * np.sin(x / 7.0) * np.cos(y / 11.0)
    g = np.where(x + 2 * y > W, 200.0, 40.0) + 10 * rng.standard_normal((H, W))
    b = 60 + 150 * (x / max(W - 1, 1))
    return np.clip(np.stack([r, g, b], -1), 0, 255).round().astype(np.uint8)


@pytest.mark.parametrize("sampling", [(1, 1), (2, 2), (2, 1), (1, 2)])
@pytest.mark.parametrize("W,H", [(64, 48), (37, 23)])
def test_progressive_equals_sequential(cg, sampling, W, H):
    img = _photo(W, H)
    q = np.clip(np.add.outer(np.arange(8), np.arange(8)) * 3 + 2, 1, 255)
    comps, tables = jf.coefficients(img, sampling, q)
    seq = decode(cg, jf.encode(W, H, comps, tables))
    prog = decode(cg, jf.encode(W, H, comps, tables, progressive=True))
    assert np.array_equal(seq, prog)
    for restart in (1, 3):
        assert np.array_equal(decode(cg, jf.encode(W, H, comps, tables, restart=restart)), seq)
        assert np.array_equal(decode(cg, jf.encode(W, H, comps, tables, progressive=True, restart=restart)), seq)


def test_near_lossless_round_trip(cg):
    """All-ones quantisation, no subsampling: the image comes back within the
    rounding of the DCT coefficients and of the colour conversion."""
    img = _photo(48, 32, seed=7)
    comps, tables = jf.coefficients(img)
    got = decode(cg, jf.encode(48, 32, comps, tables))[..., :3].astype(int)
    d = np.abs(got - img.astype(int))
    assert d.max() <= 3 and d.mean() < 0.6, (d.max(), d.mean())


@pytest.mark.parametrize("sampling", [(2, 2), (2, 1), (3, 1)])
def test_subsampled_chroma_is_upsampled(cg, sampling):
    """A smooth image with chroma at half (or a third) resolution decodes close
    to the image; flat colour decodes exactly, whatever the factors."""
    y, x = np.mgrid[0:32, 0:48]
    img = np.stack([60 + 2 * x, 200 - 3 * y, 90 + x + y], -1).clip(0, 255).astype(np.uint8)
    comps, tables = jf.coefficients(img, sampling)
    got = decode(cg, jf.encode(48, 32, comps, tables))[..., :3].astype(int)
    d = np.abs(got - img.astype(int))
    assert d.mean() < 1.5 and d.max() <= 8, (d.mean(), d.max())
    flat = np.zeros((32, 48, 3), np.uint8) + np.array([10, 150, 240], np.uint8)
    comps, tables = jf.coefficients(flat, sampling)
    got = decode(cg, jf.encode(48, 32, comps, tables))[..., :3].astype(int)
    assert np.abs(got - flat.astype(int)).max() <= 1


def test_upsampling_weights(cg):
    """2:1 horizontal chroma: a Cr ramp of one step per chroma sample becomes,
    at full resolution, the 3/4 - 1/4 blend of the nearest two samples: with
    Y = 128 and Cb = 128 the red channel follows 128 + 1.402 (Cr - 128)."""
    W, H = 32, 8
    comps, tables = jf.coefficients(np.full((H, W), 128, np.uint8))
    y = comps[0]
    y["h"], y["v"] = 2, 1
    cr_plane = np.tile(128 + 8 * np.arange(16, dtype=np.float64), (8, 1))  # 16 chroma samples across
    cb = {"h": 1, "v": 1, "tq": 0, "bw": 2, "bh": 1,
          "blocks": np.zeros((1, 2, 8, 8), np.int64)}
    cr_blocks = np.zeros((1, 2, 8, 8), np.int64)
    for bx in range(2):
        f = jf.fdct(cr_plane[:, bx * 8:bx * 8 + 8] - 128.0)
        cr_blocks[0, bx] = np.sign(f) * np.floor(np.abs(f) + 0.5)
    cr = {"h": 1, "v": 1, "tq": 0, "bw": 2, "bh": 1, "blocks": cr_blocks}
    got = decode(cg, jf.encode(W, H, [y, cb, cr], tables))
    # the decoded chroma samples, from the exact inverse DCT of what was sent
    samples = np.concatenate([np.floor(jf.idct(cr_blocks[0, bx].astype(float)) + 128.5) for bx in range(2)], axis=1)[0]
    for xo in range(2, W - 2):
        s = xo // 2
        n = s - 1 if xo % 2 == 0 else s + 1
        crv = 0.75 * samples[s] + 0.25 * samples[n]
        expect = np.clip(np.floor(128 + 1.402 * (crv - 128) + 0.5), 0, 255)
        assert abs(int(got[4, xo, 0]) - expect) <= 0, (xo, got[4, xo, 0], expect)


@pytest.mark.parametrize("mutate,msg", [
    (lambda d: d[:40], "truncated|without image"),
    (lambda d: d.replace(b"\xff\xc0", b"\xff\xc3", 1), "lossless"),
    (lambda d: d.replace(b"\xff\xc0", b"\xff\xc9", 1), "arithmetic"),
    (lambda d: b"\xff\xd8" + d[21:], "quantisation"),  # the tables' segment cut: its data skipped
    (lambda d: d.replace(b"\xff\xda", b"\xff\xdf", 1), "without image data|truncated|length"),
])
def test_malformed_and_unsupported(cg, mutate, msg):
    comps, tables = jf.coefficients(_photo(16, 16))
    data = jf.encode(16, 16, comps, tables)
    with pytest.raises(ValueError, match=msg):
        decode(cg, mutate(data))


def test_four_components_refused(cg):
    data = bytearray(jf.encode(8, 8, *jf.coefficients(_photo(8, 8))))
    i = data.index(b"\xff\xc0")
    data[i + 9] = 4  # component count in the frame header
    with pytest.raises(ValueError, match="4 components"):
        decode(cg, bytes(data))


def test_gltf_with_jpeg_textures_loads(cg, tmp_path):
    """The glTF test scene with its PNG images replaced by JPEG ones (base
    colour as a data URI, the others as well): LoadScene accepts it and keeps
    exactly the decoder's texels."""
    p = gf.write(tmp_path, "datauri")
    doc = json.loads(p.read_text())
    jpegs = []
    for i, im in enumerate(doc["images"]):
        img = _photo(24 + 8 * i, 16 + 8 * i, seed=i)
        comps, tables = jf.coefficients(img, (2, 2) if i % 2 else (1, 1),
                                        np.clip(np.add.outer(np.arange(8), np.arange(8)) + 1, 1, 255))
        data = jf.encode(img.shape[1], img.shape[0], comps, tables, progressive=bool(i % 2))
        jpegs.append(data)
        doc["images"][i] = {"uri": "data:image/jpeg;base64," + base64.b64encode(data).decode()}
    p.write_text(json.dumps(doc))
    s = cg.LoadScene(str(p))
    tex = s.describe()["textures"]
    assert len(tex) >= len(jpegs)
    by_image = {}
    for t, src in zip(tex, [t.get("source") for t in doc["textures"]]):
        by_image.setdefault(src, t)
    for i, data in enumerate(jpegs):
        if i in by_image:
            assert np.array_equal(by_image[i]["rgba8"], decode(cg, data))
