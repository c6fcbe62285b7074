"""GPU: the render schedule.  From its second render of one size on, an
environment takes the work items (sample-stream mode: a tile's samples in
chunks; reference order: tiles) longest first by the previous render's
durations; that decides which wave runs an item, never its result.  Bar:
every frame of a progressive sequence equals, bit for bit (XYZ history,
carried seed, path records), the same frame rendered by a fresh environment
from the same history in image order, including after a render of another
size has reset the schedule and with the schedule switched off
(cg_env_set_schedule)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fresh_frame(cg, conf, sd, cam, it, hist, rng_mode):
    with cg.DeviceEnv(conf) as env:
        ds = env.upload(sd)
        env.set_rng_mode(rng_mode)
        env.set_path_records(True)
        if hist is not None:
            env.write_history(hist)
        env.view(ds, cam, it)
        return env.read_history(), env.read_path_records()


@pytest.mark.parametrize("rng_mode", [0, 1])
@pytest.mark.parametrize("name,params,w,h,spp,depth", [
    ("c1", [0.25], 160, 96, 16, 8),
    ("spheres", [8], 96, 64, 40, 5),
])
def test_scheduled_frames_equal_image_order(cg, name, params, w, h, spp, depth, rng_mode):
    s = cg.HostScene.generate(name, seed=5, params=params)
    sd = s.scene_data()
    cam = s.describe()["camera"]
    conf = cg.Conf.make(w, h, spp, depth)
    with cg.DeviceEnv(conf) as env:
        ds = env.upload(sd)
        env.set_rng_mode(rng_mode)
        env.set_path_records(True)
        prev = None
        for it in range(5):
            if it == 4:  # the schedule switched off: image order
                env.set_schedule(False)
            if it == 2:  # a render of another size in this env resets the schedule
                job = cg.RenderJob(env, ds, cam, cg.Conf.make(w // 2, h // 2, spp, depth), rng_mode=rng_mode)
                job.view_alone()
                env.sync()
                job.free()
            env.view(ds, cam, it)
            hist = env.read_history()
            ph, rc = env.read_path_records()
            ref_hist, (ref_ph, ref_rc) = _fresh_frame(cg, conf, sd, cam, it, prev, rng_mode)
            print(f"{name} rng mode {rng_mode} frame {it}: history equal {np.array_equal(hist, ref_hist)}, "
                  f"paths equal {np.array_equal(ph, ref_ph)}", flush=True)
            assert np.array_equal(hist.view(np.uint8), ref_hist.view(np.uint8))
            assert np.array_equal(ph, ref_ph)
            assert np.array_equal(rc, ref_rc)
            prev = hist
