"""Host-side checks (no GPU): the C-ABI libraries load and export every
declared symbol, the crystal::graphics API's error behaviour mirrors the
reference's, and the product's scene preparation equals the oracle's
restatement of the reference's (src/pathtracing/*.h) bit for bit."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def _declared(header: Path) -> list[str]:
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cg_[a-z_0-9]+)\s*\(", text)))


def test_device_library_exports_every_declared_symbol(cg):
    from crystalgraphics_amd import native
    lib = native.device()
    names = _declared(ROOT / "include" / "cg_device.h")
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_host_library_exports(cg):
    from crystalgraphics_amd import native
    h = native.host()
    for n in ("cgh_scene_generate", "cgh_scene_data_build", "cgh_env_init", "cgh_view", "cgh_view_to_buffer",
              "cgh_load_scene", "cgh_env_term", "cgh_submit_jobs", "cgh_server_create"):
        assert hasattr(h, n)


def test_scene_records_match_header_sizes(cg):
    # sizes asserted against include/cg_scene.h's reference layouts
    assert cg.VERTEX.itemsize == 48 and cg.MATERIAL.itemsize == 432
    assert cg.TLAS_NODE.itemsize == 32 and cg.INSTANCE.itemsize == 80 and cg.BLAS_NODE.itemsize == 32
    assert cg.INDEX.itemsize == 16 and cg.EMISSIVE.itemsize == 80 and cg.ALIAS.itemsize == 8


def _has_gpu(cg) -> bool:
    try:
        e = cg.DeviceEnv(cg.Conf.make(8, 8))
    except cg.NativeError:
        return False
    e.close()
    return True


def test_env_create_without_device_fails_cleanly(cg):
    if _has_gpu(cg):
        pytest.skip("a HIP device is present")
    with pytest.raises(cg.NativeError) as ei:
        cg.DeviceEnv(cg.Conf.make(8, 8))
    assert ei.value.code == 2 and "HIP" in ei.value.message


def test_env_create_rejects_bad_conf(cg):
    from crystalgraphics_amd import native
    h = C.c_void_p()
    rc = native.device().cg_env_create(C.byref(cg.Conf.make(0, 8)), 0, C.byref(h))
    assert rc == 1 and b"render size" in native.device().cg_last_error()


def test_null_arguments_are_rejected(cg):
    from crystalgraphics_amd import native
    d = native.device()
    assert d.cg_env_destroy(None) == 1
    assert d.cg_env_sync(None) == 1
    assert d.cg_view(None, None, None, 0) == 1
    assert d.cg_trace_closest(None, None, None, 4, None, None) == 1
    assert d.cg_stream_copy(None, None, None, 16) == 1
    assert d.cg_scene_refit(None, None, None) == 1
    h = native.host()
    assert h.cgh_submit_jobs(None, 1, None, None, None, 1, None, None) == 1
    assert h.cgh_server_submit(None, 0, None, None, None) != 0
    assert h.cgh_server_next(None, 0.0, None, None, 0) != 0


def test_api_error_paths(cg):
    """View before EnvInit, EnvTerm twice, an unknown scene: the reference's
    messages (src/view.cpp:13-14, src/wgpu/environment.cpp:127-128)."""
    assert not cg.EnvStatus()
    s = cg.HostScene.generate("quad")
    with pytest.raises(cg.NativeError, match="Environment not initialized"):
        cg.View(s, s.describe()["camera"])
    with pytest.raises(cg.NativeError, match="Environment not initialized"):
        cg.SubmitJobs([s], [(0, s.describe()["camera"], cg.Conf.make(8, 8, 1, 2))])
    with pytest.raises(cg.NativeError, match="already terminated"):
        cg.EnvTerm()
    with pytest.raises(cg.NativeError, match="not supported"):
        cg.LoadScene("/nonexistent/scene.obj")
    with pytest.raises(cg.NativeError, match="cannot read"):
        cg.LoadScene("/nonexistent/scene.gltf")
    with pytest.raises(cg.NativeError, match="unknown procedural scene"):
        cg.LoadScene("procedural:nope")
    sc = cg.LoadScene("procedural:spheres?seed=4")
    assert sc.describe()["prims"].shape[0] == 10


@pytest.mark.parametrize("name,params", [("quad", []), ("spheres", [8]), ("c0", []), ("c1", [0.25]),
                                         ("spheres", [8, 1, 64]), ("c1", [0.25, 1, 128])])
def test_scene_prep_matches_oracle(cg, oracle, name, params):
    s = cg.HostScene.generate(name, seed=7, params=params)
    a = s.scene_data(share_blas=False, threads=8).arrays()
    b = oracle.OracleScene.build(s.describe()).arrays()
    for k in a:
        assert a[k].tobytes() == b[k].tobytes(), k


def test_c0_matches_published_survey_counts(cg, oracle):
    s = cg.HostScene.generate("c0", seed=1)
    a = s.scene_data(share_blas=False).arrays()
    assert len(s.describe()["indices"]) // 3 == 20484
    assert len(a["blas_nodes"]) == 24514 and len(a["tlas_nodes"]) == 131 and len(a["instances"]) == 66


def test_parallel_blas_build_is_identical(cg):
    s = cg.HostScene.generate("c2", seed=3, params=[300000])
    a = s.scene_data(threads=1).arrays()
    b = s.scene_data(threads=16).arrays()
    for k in ("blas_nodes", "indices", "tri_ids"):
        assert a[k].tobytes() == b[k].tobytes(), k


def test_shared_blas_is_the_same_tree(cg):
    s = cg.HostScene.generate("c1", seed=2, params=[0.25])
    shared = s.scene_data(share_blas=True).arrays()
    full = s.scene_data(share_blas=False).arrays()
    assert len(shared["blas_nodes"]) < len(full["blas_nodes"])
    roots = set(int(r) for r in shared["instances"]["blas_root_idx"][:16])
    assert len(roots) == 1
    r_full = full["instances"]["blas_root_idx"]
    n = int(r_full[1] - r_full[0])
    t0 = full["blas_nodes"][r_full[0]:r_full[0] + n]
    t1 = full["blas_nodes"][r_full[1]:r_full[1] + n]
    # identical boxes and counts; child indices offset by the copy's base
    assert np.array_equal(t0["lb"], t1["lb"]) and np.array_equal(t0["ub"], t1["ub"])
    assert np.array_equal(t0["triangle_count"], t1["triangle_count"])


def test_generated_workloads_have_the_stated_shape(cg):
    c1 = cg.HostScene.generate("c1", seed=1).describe()
    prims = c1["prims"]
    assert len(prims) == 16 + 4096
    assert prims["index_count"][0] // 3 == 131072
    assert np.all(prims["index_count"][16:] == 3)
    strengths = c1["materials"]["emissive_strength"][prims["material_idx"][16:]]
    assert strengths.min() >= 1.0 and strengths.max() <= 100.0
    c2 = cg.HostScene.generate("c2", seed=1, params=[100000]).describe()
    v = c2["vertices"]["pos"]
    assert len(c2["indices"]) // 3 == 100000 and v.min() >= -0.02 and v.max() <= 1.02


def test_ray_batch_halves(cg):
    s = cg.HostScene.generate("c2", seed=1, params=[20000])
    r = s.rays(4096, seed=1)
    coherent, incoherent = r[:2048], r[2048:]
    assert np.allclose(coherent["org"], coherent["org"][0])
    assert np.allclose(np.linalg.norm(r["dir"], axis=1), 1, atol=1e-5)
    assert np.ptp(incoherent["org"], axis=0).min() > 0.1


@pytest.mark.parametrize("name,params,accel", [("c0", [], 1), ("c0", [], 0), ("c1", [0.25], 1), ("spheres", [8], 1),
                                               ("c2", [200000], 1), ("c0", [], 2), ("c1", [0.25], 2),
                                               ("spheres", [8], 2), ("c2", [200000], 2), ("quad", [], 2)])
def test_quantised_nodes_contain_exact_boxes(cg, name, params, accel):
    """The device layout's 128-B quantised 8-wide nodes decode (with the kernels' own
    float operations) to boxes that contain the exact child boxes, so the
    traversal visits a superset of the exact tree's nodes; every triangle
    slot sits in exactly one leaf.  Accel 2 runs the device linear-BVH build's
    per-thread steps on the host."""
    from crystalgraphics_amd import native
    s = cg.HostScene.generate(name, seed=2, params=params)
    sd = s.scene_data()
    n, depth = C.c_uint64(), C.c_uint32()
    rc = native.device().cg_layout_check(C.byref(sd.c), accel, C.byref(n), C.byref(depth))
    assert rc == 0, native.device().cg_last_error()
    assert (n.value > 0 or name == "quad") and 1 <= depth.value <= 128  # quad: one instance, one leaf


def test_layout_of_huge_coordinates(cg):
    """Triangles spread over 1e20 units: the collapse's surface-area costs
    overflow to infinity in float, and the build must still produce a valid
    8-wide tree (the containment and one-leaf-per-slot checks of
    cg_layout_check)."""
    from crystalgraphics_amd import native
    rng = np.random.default_rng(4)
    n = 300
    base = rng.uniform(-1e20, 1e20, (n, 1, 3))
    pos = (base + rng.uniform(-1e18, 1e18, (n, 3, 3))).reshape(-1, 3)
    v = np.zeros(len(pos), cg.VERTEX)
    v["pos"] = pos
    v["nrm"] = [0, 0, 1]
    m = np.zeros(1, cg.MATERIAL)
    m["base_color_factor"] = [0.5, 0.5, 0.5, 1]
    m["roughness_factor"] = 1
    m["ior"] = 1.5
    m["tex"]["index"] = -1
    p = np.zeros(1, cg.PRIMITIVE)
    p["vertex_count"] = len(pos)
    p["index_count"] = len(pos)
    p["world"] = np.eye(4, dtype=np.float32).reshape(-1)
    p["inv_world"] = np.eye(4, dtype=np.float32).reshape(-1)
    s = cg.HostScene.from_arrays(v, np.arange(len(pos), dtype=np.uint32), m, p, tag=93)
    sd = s.scene_data()
    nodes, depth = C.c_uint64(), C.c_uint32()
    rc = native.device().cg_layout_check(C.byref(sd.c), 1, C.byref(nodes), C.byref(depth))
    assert rc == 0, native.device().cg_last_error()
    assert nodes.value > 0 and 1 <= depth.value <= 128
