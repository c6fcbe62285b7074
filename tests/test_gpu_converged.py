"""GPU: converged images.  A 4096-spp render on the GPU against the oracle's
4096-spp render of the same scene and camera: image relative MSE
sum((gpu - cpu)^2) / sum(cpu^2) over all pixels and XYZ channels must be
below 1e-3 (bit-identical paths make it ~0; a diverged path changes one
sample of one pixel)."""
import time

import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]


@pytest.mark.parametrize("name,params,w,h,depth", [("spheres", [8], 16, 12, 3), ("c0", [], 32, 18, 2)])
def test_converged_image_matches_cpu(cg, oracle, name, params, w, h, depth):
    s = cg.HostScene.generate(name, seed=9, params=params)
    sd = s.scene_data()
    cam = s.describe()["camera"]
    if name == "c0":
        cam = cg.Camera.make(cam.position[:], cam.direction[:], (0.5 * w / h, 0.5))
    conf = cg.Conf.make(w, h, 4096, depth)
    with cg.DeviceEnv(conf) as env:
        ds = env.upload(sd)
        env.view(ds, cam, 0)
        gpu = env.read_history()["xyz"].astype(np.float64)
    t0 = time.time()
    orc = oracle.OracleScene.wrap(sd.c)
    cpu = orc.render(conf, cam, threads=16)["history"]["xyz"].astype(np.float64)
    orc.close()
    rmse = ((gpu - cpu) ** 2).sum() / max((cpu ** 2).sum(), 1e-30)
    print(f"{name} {w}x{h} 4096 spp D{depth}: relative MSE {rmse:.3e} (oracle {time.time() - t0:.1f} s)")
    assert np.isfinite(gpu).all() and (cpu ** 2).sum() > 0
    assert rmse < 1e-3
