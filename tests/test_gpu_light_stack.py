"""GPU: the light pdf of next-event estimation (pdf/emission.wesl:120-149)
where a direction passes through many emitters.  Ten large emissive
triangles of different strengths are stacked above a diffuse floor, so every
upward direction from the floor crosses all ten; the light pdf then sums ten
terms, found in the emitter BVH out of index order, and must add them in
ascending emitter index as the reference's loop does (with 24 stacked
emitters the device takes its loop over every emitter instead).  Bar: per pixel against
the oracle (which sums over every emitter in index order), in both RNG
modes -- paths identical on >= 99 % of pixels and XYZ bit-exact on those."""
import numpy as np
import pytest

from helpers import pixel_rel_err

pytestmark = pytest.mark.gpu

N_LIGHTS = 10


def _stacked_scene(cg, n_lights=N_LIGHTS):
    pos = [[-2, -2, 0], [2, -2, 0], [2, 2, 0], [-2, 2, 0]]
    idx = [0, 1, 2, 0, 2, 3]
    for k in range(n_lights):
        z = 1.0 + 0.9 * k / n_lights
        s = 0.1 * k
        base = len(pos)
        pos += [[-7 + s, -6, z], [7, -6 - s, z], [s, 9, z]]
        idx += [base, base + 1, base + 2]
    v = np.zeros(len(pos), cg.VERTEX)
    v["pos"] = pos
    v["nrm"] = [[0, 0, 1]] * 4 + [[0, 0, -1]] * (len(pos) - 4)
    m = np.zeros(1 + n_lights, cg.MATERIAL)
    m["base_color_factor"] = [[0.7, 0.6, 0.5, 1]] + [[0, 0, 0, 1]] * n_lights
    m["metallic_factor"] = 0
    m["roughness_factor"] = 1
    m["ior"] = 1.5
    m["tex"]["index"] = -1
    m["emissive_factor"][1:] = [[1.0, 0.8, 0.6]] * n_lights
    m["emissive_strength"] = [1.0] + [0.5 + 0.7 * k for k in range(n_lights)]
    p = np.zeros(1 + n_lights, cg.PRIMITIVE)
    p["vertex_offset"] = [0] + [4 + 3 * k for k in range(n_lights)]
    p["vertex_count"] = [4] + [3] * n_lights
    p["index_offset"] = [0] + [6 + 3 * k for k in range(n_lights)]
    p["index_count"] = [6] + [3] * n_lights
    p["material_idx"] = list(range(1 + n_lights))
    p["world"] = np.eye(4, dtype=np.float32).reshape(-1)
    p["inv_world"] = np.eye(4, dtype=np.float32).reshape(-1)
    env = np.zeros((1, 1, 4), np.float32)
    env[0, 0] = [0.05, 0.05, 0.08, 1.0]
    return cg.HostScene.from_arrays(v, np.array(idx, np.uint32), m, p, env=env, tag=91)


@pytest.mark.parametrize("n_lights", [N_LIGHTS, 24])
@pytest.mark.parametrize("rng_mode", [0, 1])
def test_light_pdf_through_stacked_emitters(cg, oracle, rng_mode, n_lights):
    """Ten emitters: the BVH walk once per emitter crossed; 24: beyond the
    walks' bound, the loop over every emitter."""
    s = _stacked_scene(cg, n_lights)
    sd = s.scene_data()
    d = np.array([0.0, 1.0, -0.45])
    cam = cg.Camera.make([0.0, -1.9, 0.6], d / np.linalg.norm(d))
    conf = cg.Conf.make(48, 32, 4, 3)
    orc = oracle.OracleScene.wrap(sd.c)
    with cg.DeviceEnv(conf) as env:
        ds = env.upload(sd)
        env.set_rng_mode(rng_mode)
        env.set_path_records(True)
        env.view(ds, cam, 0)
        hist = env.read_history()
        ph, _ = env.read_path_records()
    ref = orc.render(conf, cam, rng_mode=rng_mode)
    same = ph == ref["path_hash"]
    err = pixel_rel_err(hist["xyz"], ref["history"]["xyz"])
    lit = np.abs(ref["history"]["xyz"]).max(axis=-1) > 0
    print(f"stacked emitters, rng mode {rng_mode}: {same.mean() * 100:.2f}% identical paths, "
          f"{lit.mean() * 100:.1f}% lit pixels, max rel err {err[same].max():.2e}", flush=True)
    assert lit.mean() > 0.9
    assert same.mean() >= 0.99
    assert err[same].max() == 0.0

