"""GPU: the crystal::graphics facade (EnvInit, LoadScene, View, ViewToBuffer,
EnvPoll, EnvTerm; include/CrystalGraphics/api.h in the reference) with its
progressive accumulation: the frame history keeps accumulating while the
scene and camera stay the same and restarts when either changes
(src/wgpu/frame_history.h:12-47, compute.wesl:41-56).  Every frame is checked
against the oracle's render of the same iteration, to the per-pixel bar."""
import numpy as np
import pytest

from helpers import pixel_rel_err

pytestmark = pytest.mark.gpu


def _check(out, ref, label):
    err = pixel_rel_err(out, ref["history"]["xyz"])
    ok = err <= 1e-4
    print(f"{label}: {ok.mean() * 100:.2f}% of pixels within 1e-4")
    assert ok.mean() >= 0.99, label


def test_facade_accumulates_and_restarts(cg, oracle):
    conf = cg.Conf.make(48, 32, 2, 4)
    s = cg.LoadScene("procedural:spheres?seed=4")
    sd = s.scene_data()
    orc = oracle.OracleScene.wrap(sd.c)
    cam = s.describe()["camera"]
    p = np.array(cam.position[:], np.float64)
    cam2 = cg.Camera.make(p + np.array([0.3, 0.0, 0.1]), np.array(cam.direction[:], np.float64), cam.viewport[:])
    out = np.zeros((32, 48, 3), np.float32)
    cg.EnvInit(conf)
    try:
        hist = None
        for it in range(3):  # same scene and camera: iterations 0, 1, 2 of one accumulation
            cg.ViewToBuffer(s, cam, out)
            ref = orc.render(conf, cam, iter_count=it, history=hist)
            hist = ref["history"]
            _check(out, ref, f"frame {it}")
        cg.ViewToBuffer(s, cam2, out)  # camera moved: a new accumulation
        _check(out, orc.render(conf, cam2, iter_count=0), "camera change")
        q = cg.LoadScene("procedural:quad")
        qsd = q.scene_data()
        qcam = q.describe()["camera"]
        cg.ViewToBuffer(q, qcam, out)  # another scene: a new accumulation
        _check(out, oracle.OracleScene.wrap(qsd.c).render(conf, qcam, iter_count=0), "scene change")
        cg.ViewToBuffer(s, cam2, out)  # back to the first scene: restarts again
        _check(out, orc.render(conf, cam2, iter_count=0), "scene change back")
        cg.View(s, cam2)  # presentation path: accumulates without a read-back
        cg.EnvPoll()
        cg.EnvPresent()
        cg.ViewToBuffer(s, cam2, out)
        r1 = orc.render(conf, cam2, iter_count=0)
        r2 = orc.render(conf, cam2, iter_count=1, history=r1["history"])
        _check(out, orc.render(conf, cam2, iter_count=2, history=r2["history"]), "after View")
    finally:
        cg.EnvTerm()
    assert not cg.EnvStatus()


def test_gltf_scene_renders_like_the_oracle(cg, oracle, tmp_path):
    """A glTF file (textures, glass, an emissive textured material, a node
    hierarchy) loaded by LoadScene renders on the device like the oracle:
    paths identical on >= 99 % of pixels, XYZ within 1e-4 relative on those."""
    import gltf_files as gf

    s = cg.LoadScene(str(gf.write(tmp_path, "glb")))
    sd = s.scene_data()
    cam = s.describe()["camera"]
    conf = cg.Conf.make(96, 54, 4, 5)
    with cg.DeviceEnv(conf) as env:
        ds = env.upload(sd)
        env.set_path_records(True)
        env.view(ds, cam, 0)
        hist = env.read_history()
        phash, rcount = env.read_path_records()
    ref = oracle.OracleScene.wrap(sd.c).render(conf, cam)
    same = phash == ref["path_hash"]
    err = pixel_rel_err(hist["xyz"], ref["history"]["xyz"])
    print(f"glTF scene: paths identical {same.mean() * 100:.2f}%, max rel err {err[same].max():.3e}, "
          f"lit pixels {(np.abs(ref['history']['xyz']).max(axis=-1) > 0).mean() * 100:.1f}%")
    assert same.mean() >= 0.99 and err[same].max() <= 1e-4
    assert np.array_equal(rcount[same], ref["ray_count"][same])


def test_gltf_environment_renders_like_the_oracle(cg, oracle, tmp_path):
    """A glTF file whose extras name a Radiance HDR sky: LoadScene decodes it
    into the scene's environment texture, and the device render equals the
    oracle's on the same scene data (paths >= 99 %, XYZ within 1e-4)."""
    import json

    import gltf_files as gf
    rng = np.random.default_rng(12)
    sky = rng.integers(0, 256, (32, 64, 4)).astype(np.uint8)
    sky[..., 3] = rng.integers(120, 140, (32, 64))
    (tmp_path / "sky.hdr").write_bytes(gf.hdr_file(sky, rle=True))
    p = gf.write(tmp_path, "datauri")
    doc = json.loads(p.read_text())
    doc["extras"] = {"environment_hdr": "sky.hdr"}
    p.write_text(json.dumps(doc))
    s = cg.LoadScene(str(p))
    sd = s.scene_data()
    assert s.describe()["env"].shape == (32, 64, 4)
    cam = s.describe()["camera"]
    conf = cg.Conf.make(80, 45, 4, 5)
    with cg.DeviceEnv(conf) as env:
        ds = env.upload(sd)
        env.set_path_records(True)
        env.view(ds, cam, 0)
        hist = env.read_history()
        phash, _ = env.read_path_records()
    ref = oracle.OracleScene.wrap(sd.c).render(conf, cam)
    same = phash == ref["path_hash"]
    err = pixel_rel_err(hist["xyz"], ref["history"]["xyz"])
    print(f"glTF scene with HDR sky: paths identical {same.mean() * 100:.2f}%, max rel err {err[same].max():.3e}")
    assert same.mean() >= 0.99 and err[same].max() <= 1e-4


def test_submit_jobs_equal_single_renders(cg):
    """SubmitJobs (many requests, two scenes, one launch) returns for every
    request exactly the image of rendering it alone with the plain kernel."""
    from crystalgraphics_amd import workloads
    scenes = [cg.HostScene.generate("spheres", seed=4, params=[6]),
              cg.HostScene.generate("c1", seed=3, params=[0.125, 1, 128])]
    rng = np.random.default_rng(9)
    reqs = []
    for k in range(7):
        w, h = [(32, 24), (48, 40), (17, 11), (64, 36)][k % 4]
        pos, d = workloads.orbit_camera(rng, 3.0 if k % 2 else 12.0)
        reqs.append((k % 2, cg.Camera.make(pos, d, (0.5 * w / h, 0.5)), cg.Conf.make(w, h, [1, 2, 4][k % 3], [2, 5, 8][k % 3])))
    cg.EnvInit(cg.Conf.make(64, 64, 1, 2))
    try:
        outs, secs = cg.SubmitJobs(scenes, reqs)
    finally:
        cg.EnvTerm()
    assert secs > 0
    with cg.DeviceEnv(cg.Conf.make(8, 8)) as env:
        sds = [s.scene_data() for s in scenes]
        dss = [env.upload(sd) for sd in sds]
        for (si, cam, conf), got in zip(reqs, outs):
            rj = cg.RenderJob(env, dss[si], cam, conf)
            rj.view_alone()
            ref = rj.read_history()["xyz"].reshape(conf.render_height, conf.render_width, 3)
            assert got.tobytes() == ref.tobytes()
            rj.free()


def test_stack_overflow_is_reported(cg, monkeypatch):
    """A traversal whose stack outgrows what the launch provides sets a device
    error flag that cg_env_sync reports (and clears).  Forced from the host:
    CG_TEST_STACK_CAP=0 gives the launch no overflow entries past the LDS part,
    on a scene whose stack bound is deeper; pushes beyond are dropped, so every
    kernel still finishes."""
    s = cg.HostScene.generate("c0", seed=11)
    sd = s.scene_data()
    rays = s.rays(1 << 14, seed=3)
    monkeypatch.setenv("CG_TEST_STACK_CAP", "0")
    env = cg.DeviceEnv(cg.Conf.make(16, 16, 1, 1))
    monkeypatch.delenv("CG_TEST_STACK_CAP")
    try:
        env.set_accel(0)
        ds = env.upload(sd)
        assert ds.footprint()[1] > 16
        env.trace_closest(ds, rays)
        with pytest.raises(cg.NativeError, match="stack overflow"):
            env.sync()
        env.sync()  # reported once, then clear
    finally:
        env.close()
    ok = cg.DeviceEnv(cg.Conf.make(16, 16, 1, 1))
    try:
        ok.set_accel(0)
        ds = ok.upload(sd)
        ok.trace_closest(ds, rays)
        ok.sync()  # the stack bound holds: no report
    finally:
        ok.close()
