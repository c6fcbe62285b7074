"""GPU: the material texture path (bilinear repeat lookups of the base colour,
metallic-roughness, transmission and emissive textures, through either uv
set) against the oracle, and the white-texture identity on the device."""
import numpy as np
import pytest

from helpers import pixel_rel_err

pytestmark = pytest.mark.gpu


def test_white_textures_change_nothing_on_device(cg):
    """All-255 textures on every slot of every material: the device renders
    the scene bit for bit like the untextured one (history, rgba8 and path
    records)."""
    plain = cg.HostScene.generate("spheres", seed=9, params=(6,))
    d = plain.describe()
    mats = d["materials"].copy()
    for i in range(1, len(mats)):
        for slot, coord in ((0, 0), (1, 1), (15, 0), (4, 1)):
            mats[i]["tex"][slot]["index"] = slot % 4
            mats[i]["tex"][slot]["tex_coord"] = coord
    white = [{"rgba8": np.full((h, w, 4), 255, np.uint8)} for h, w in ((4, 4), (3, 5), (1, 1), (8, 2))]
    textured = cg.HostScene.from_arrays(d["vertices"], d["indices"], mats, d["prims"], textures=white, tag=77)
    conf = cg.Conf.make(96, 64, 4, 4)
    cam = d["camera"]
    outs = []
    for sc in (plain, textured):
        sd = sc.scene_data()
        with cg.DeviceEnv(conf) as env:
            ds = env.upload(sd)
            env.set_path_records(True)
            env.view(ds, cam, 0)
            outs.append((env.read_history().tobytes(), env.read_rgba8().tobytes(), env.read_path_records()))
    assert outs[0][0] == outs[1][0] and outs[0][1] == outs[1][1]
    assert np.array_equal(outs[0][2][0], outs[1][2][0]) and np.array_equal(outs[0][2][1], outs[1][2][1])


def test_full_size_textures_rows(cg, oracle):
    """Four 2048 x 2048 textures (64 MiB of texels, the serving scenes'
    size) on a 1280 x 720 frame of the 16-sphere scene: three bands of rows
    against the oracle, to the per-pixel bar (paths identical on >= 99 % of
    pixels, XYZ within 1e-4 relative on those)."""
    s = cg.HostScene.generate("c1", seed=3, params=[0.25, 1, 2048])
    sd = s.scene_data()
    assert sd.arrays()["tex"].shape == (4, 2048, 2048, 4)
    cam = s.describe()["camera"]
    conf = cg.Conf.make(1280, 720, 1, 6)
    orc = oracle.OracleScene.wrap(sd.c)
    with cg.DeviceEnv(conf) as env:
        ds = env.upload(sd)
        env.set_path_records(True)
        env.view(ds, cam, 0)
        hist = env.read_history()
        phash, rcount = env.read_path_records()
    for y0 in (96, 360, 600):
        ref = orc.render(conf, cam, rows=(y0, y0 + 8), threads=16)
        sl = slice(y0, y0 + 8)
        same = phash[sl] == ref["path_hash"][sl]
        err = pixel_rel_err(hist["xyz"][sl], ref["history"]["xyz"][sl])
        print(f"rows {y0}-{y0 + 8}: paths identical {same.mean()*100:.2f}%, max rel err {err[same].max():.3e}")
        assert same.mean() >= 0.99
        assert err[same].max() <= 1e-4
        assert np.array_equal(rcount[sl][same], ref["ray_count"][sl][same])
