"""GPU: refitting device-built BVHs (cg_scene_refit).  A scene uploaded with
every BLAS built on the device takes new data with the same triangles --
moved vertices, moved instances -- and its closest hits and frames equal the
oracle's on the new data; refitting with unchanged data changes nothing."""
import numpy as np
import pytest

from helpers import compare_hits, pixel_rel_err

pytestmark = pytest.mark.gpu


def _moved_vertices(cg, s, tag):
    d = s.describe()
    v = d["vertices"].copy()
    p = v["pos"].astype(np.float64)
    p = p * 1.1 + 0.02 * np.sin(7.0 * p[:, [1, 2, 0]])
    v["pos"] = p.astype(np.float32)
    return cg.HostScene.from_arrays(v, d["indices"], d["materials"], d["prims"], tag=tag)


def test_refit_moved_vertices(cg, oracle):
    s = cg.HostScene.generate("c2", seed=5, params=[300000])
    sd = s.scene_data()
    rays = s.rays(1 << 16, seed=3)
    with cg.DeviceEnv(cg.Conf.make(64, 64, 1, 1)) as env:
        env.set_accel(2)
        ds = env.upload(sd)
        env.set_accel(1)
        before = env.trace_closest(ds, rays)
        env.refit(ds, sd)  # unchanged data: nothing changes
        assert env.trace_closest(ds, rays).tobytes() == before.tobytes()
        s2 = _moved_vertices(cg, s, tag=77)
        sd2 = s2.scene_data()
        env.refit(ds, sd2)
        hits = env.trace_closest(ds, rays)
        ref = oracle.OracleScene.wrap(sd2.c).trace_closest(rays)
        id_bad, t_bad, n_tied = compare_hits(hits, ref)
        moved = (hits["inst"] != before["inst"]) | (hits["tri"] != before["tri"]) | (hits["t"] != before["t"])
        print(f"refit c2 300k: tied {n_tied}, id mismatches {id_bad}, hits that changed {moved.mean() * 100:.1f}%")
        assert id_bad == 0 and t_bad == 0 and moved.mean() > 0.5
        ds.release()


def test_refit_moved_instances_render(cg, oracle):
    s = cg.HostScene.generate("c1", seed=3, params=[0.125])
    d = s.describe()
    prims = d["prims"].copy()
    rng = np.random.default_rng(2)
    for i in range(16):  # the sphere instances move; the emitters stay
        w = prims[i]["world"].reshape(4, 4).T.astype(np.float64)
        w[:3, 3] += rng.uniform(-1.0, 1.0, 3)
        prims[i]["world"] = w.T.reshape(-1).astype(np.float32)
        prims[i]["inv_world"] = np.linalg.inv(w).T.reshape(-1).astype(np.float32)
    s2 = cg.HostScene.from_arrays(d["vertices"], d["indices"], d["materials"], prims, tag=78)
    sd, sd2 = s.scene_data(), s2.scene_data()
    conf = cg.Conf.make(128, 72, 2, 6)
    cam = d["camera"]
    with cg.DeviceEnv(conf) as env:
        env.set_accel(2)
        ds = env.upload(sd)
        env.refit(ds, sd2)
        env.set_path_records(True)
        env.view(ds, cam, 0)
        hist = env.read_history()
        phash, rcount = env.read_path_records()
        ds.release()
    ref = oracle.OracleScene.wrap(sd2.c).render(conf, cam)
    same = phash == ref["path_hash"]
    err = pixel_rel_err(hist["xyz"], ref["history"]["xyz"])
    print(f"refit c1: paths identical {same.mean() * 100:.2f}%, max rel err {err[same].max():.3e}")
    assert same.mean() >= 0.99 and err[same].max() <= 1e-4


def test_refit_rejects_other_scenes(cg):
    s = cg.HostScene.generate("spheres", seed=1, params=[4])
    other = cg.HostScene.generate("spheres", seed=1, params=[5])
    with cg.DeviceEnv(cg.Conf.make(8, 8)) as env:
        ds1 = env.upload(s.scene_data())  # the default layout: no refit
        with pytest.raises(cg.NativeError, match="CG_ACCEL_DEVICE"):
            env.refit(ds1, s.scene_data())
        env.set_accel(2)
        ds2 = env.upload(s.scene_data())
        with pytest.raises(cg.NativeError, match="other triangles"):
            env.refit(ds2, other.scene_data())
        ds1.release()
        ds2.release()
