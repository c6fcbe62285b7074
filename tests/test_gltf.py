"""LoadScene on glTF 2.0 files (the reference's src/scene.cpp:829-1003): the
same scene written as .gltf with an external buffer and images, as .gltf with
data URIs only, and as .glb loads identically, and as the independent
restatement in gltf_files.expected(); the loader's error paths; and the
loaded scene's preparation against the oracle's."""
import json

import numpy as np
import pytest

import gltf_files as gf


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    d = tmp_path_factory.mktemp("gltf")
    return {form: gf.write(d, form) for form in ("gltf", "datauri", "glb")}


@pytest.mark.parametrize("form", ["gltf", "datauri", "glb"])
def test_loads_the_restated_scene(cg, files, form):
    s = cg.LoadScene(str(files[form]))
    d = s.describe()
    exp = gf.expected()
    v = d["vertices"]
    ev = exp["vertices"]
    assert len(v) == len(ev)
    assert np.array_equal(v["pos"], ev[:, 0:3]) and np.array_equal(v["nrm"], ev[:, 3:6])
    assert np.array_equal(v["uv0"], ev[:, 6:8]) and np.array_equal(v["uv1"], ev[:, 8:10])
    assert np.array_equal(d["indices"], exp["indices"])
    prims = d["prims"]
    assert len(prims) == len(exp["prims"]) == 5  # node a: 2, node b: 1, node d: 2 (depth first)
    for p, ((vo, vc, io, ic, mat), world) in zip(prims, exp["prims"]):
        assert (p["vertex_offset"], p["vertex_count"], p["index_offset"], p["index_count"], p["material_idx"]) == \
            (vo, vc, io, ic, mat)
        w = p["world"].reshape(4, 4).T.astype(np.float64)
        assert np.abs(w - world).max() <= 1e-6
        inv = p["inv_world"].reshape(4, 4).T.astype(np.float64)
        assert np.abs(w @ inv - np.eye(4)).max() <= 1e-5
    # glTF's +y (up) is +z in the loaded scene: node d sits under a plain translation along x
    assert np.allclose(prims[3]["world"].reshape(4, 4).T[:3, :3], [[0.5, 0, 0], [0, 0, -0.5], [0, 0.5, 0]])
    texs = d["textures"]
    assert len(texs) == 4
    for t, e in zip(texs, exp["textures"]):
        assert np.array_equal(t["rgba8"], e)
    assert [(t["wrap_s"], t["wrap_t"]) for t in texs] == [(gf.CLAMP, gf.MIRROR), (gf.REPEAT, gf.REPEAT),
                                                          (gf.REPEAT, gf.REPEAT), (gf.CLAMP, gf.MIRROR)]


def test_materials(cg, files):
    """Default material at 0, then glTF material i at i + 1, with the extension
    fields the reference reads and LoadTextureInfo's scale -> strength copy."""
    m = cg.LoadScene(str(files["glb"])).describe()["materials"]
    default = cg.HostScene.generate("quad").describe()["materials"][0]
    assert len(m) == 3 and m[0].tobytes() == default.tobytes()
    a, g = m[1], m[2]
    assert np.allclose(a["base_color_factor"], [0.9, 0.8, 0.7, 1.0]) and a["metallic_factor"] == 0.0
    assert np.isclose(a["roughness_factor"], 0.6) and a["flags"] == 1 and a["alpha_mode"] == 1
    assert np.isclose(a["alpha_cutoff"], 0.25) and np.allclose(a["emissive_factor"], [0.2, 0.1, 0.05])
    t = a["tex"]
    assert (t[0]["index"], t[0]["tex_coord"], t[0]["scale"], t[0]["strength"]) == (0, 1, 1.0, 1.0)
    assert (t[1]["index"], t[4]["index"], t[15]["index"]) == (2, 1, -1)
    assert t[2]["index"] == 3 and np.isclose(t[2]["scale"], 0.7) and np.isclose(t[2]["strength"], 0.7)
    assert t[3]["index"] == 3 and t[3]["scale"] == 0.5 and t[3]["strength"] == 0.5
    assert g["emissive_strength"] == 6.0 and g["transmission_factor"] == 1.0 and g["tex"][15]["index"] == 2
    assert np.isclose(g["ior"], 1.45) and np.isclose(g["dispersion"], 0.3) and g["thickness_factor"] == 0.5
    assert np.allclose(g["attenuation_color"], [0.9, 0.95, 1.0]) and g["attenuation_distance"] == np.finfo(np.float32).max
    assert g["specular_factor"] == 0.5 and g["flags"] == 2 and g["metallic_factor"] == 0.0


def test_forms_load_identically(cg, files):
    descs = [cg.LoadScene(str(files[f])).describe() for f in ("gltf", "datauri", "glb")]
    for k in ("vertices", "indices", "materials", "prims"):
        assert descs[0][k].tobytes() == descs[1][k].tobytes() == descs[2][k].tobytes(), k


def test_scene_prep_matches_oracle(cg, oracle, files):
    s = cg.LoadScene(str(files["glb"]))
    a = s.scene_data(share_blas=False).arrays()
    b = oracle.OracleScene.build(s.describe()).arrays()
    for k in a:
        assert a[k].tobytes() == b[k].tobytes(), k
    assert len(a["emissive"]) > 0 and a["tex"].shape == (4, 5, 9, 4)


def _variant(tmp_path, mutate):
    p = gf.write(tmp_path, "datauri")
    doc = json.loads(p.read_text())
    mutate(doc)
    p.write_text(json.dumps(doc))
    return str(p)


@pytest.mark.parametrize("mutate,msg", [
    (lambda d: d["images"].__setitem__(0, {"uri": "data:image/jpeg;base64,/9j/4AAQ"}), "JPEG"),
    (lambda d: d["images"].__setitem__(0, {"uri": "data:image/webp;base64,UklGRg=="}), "not a PNG"),
    (lambda d: d["meshes"][1]["primitives"][0].__setitem__("mode", 1), "triangle"),
    (lambda d: d["meshes"][1]["primitives"][0].__setitem__("material", 7), "material"),
    (lambda d: d["nodes"][2].__setitem__("children", [9]), "node"),
    (lambda d: d.__setitem__("scenes", []), "no scene"),
    (lambda d: d["buffers"][0].__setitem__("uri", "."), "cannot read"),
    (lambda d: d["bufferViews"][0].__setitem__("byteOffset", -8), "bad bufferView"),
    (lambda d: d["accessors"][0].__setitem__("count", 2 ** 40), "bad accessor count"),
    (lambda d: d["accessors"][1].__setitem__("componentType", 1234), "componentType"),
    (lambda d: d["meshes"][0]["primitives"][0].__setitem__("material", -1), "bad primitive material"),
])
def test_error_paths(cg, tmp_path, mutate, msg):
    with pytest.raises(cg.NativeError, match=msg):
        cg.LoadScene(_variant(tmp_path, mutate))


def test_truncated_glb(cg, files, tmp_path):
    data = files["glb"].read_bytes()
    for cut in (10, 20, len(data) // 2, len(data) - 3):
        p = tmp_path / f"cut{cut}.glb"
        p.write_bytes(data[:cut])
        with pytest.raises(cg.NativeError):
            cg.LoadScene(str(p))


def test_malformed_and_missing_files(cg, tmp_path):
    bad = tmp_path / "bad.gltf"
    bad.write_text('{"asset": {"version": "2.0"}, "scenes": [')
    with pytest.raises(cg.NativeError, match="JSON"):
        cg.LoadScene(str(bad))
    with pytest.raises(cg.NativeError, match="cannot read"):
        cg.LoadScene(str(tmp_path / "missing.glb"))


def test_tag_follows_content(cg, tmp_path):
    """The scene tag (the facade's cache key) changes with the file's content,
    not only with its path and size."""
    p = gf.write(tmp_path, "datauri")
    t1 = cg.LoadScene(str(p)).describe()["tag"]
    doc = json.loads(p.read_text())
    doc["materials"][0]["pbrMetallicRoughness"]["roughnessFactor"] = 0.5  # same length as 0.6
    p.write_text(json.dumps(doc))
    t2 = cg.LoadScene(str(p)).describe()["tag"]
    assert t1 != t2
    assert cg.LoadScene(str(p)).describe()["tag"] == t2
