"""GPU: the serving form.  Requests of different scenes, sizes, spp and depth
arrive while others are in flight; each finished request must equal the same
request rendered alone with the plain kernel, bit for bit (history XYZ and
carried RNG state), and the server accounts for every request."""
import time

import numpy as np
import pytest

from crystalgraphics_amd import workloads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("engine", [0, 1])
def test_serve_results_equal_alone(cg, engine):
    scenes = [cg.HostScene.generate("spheres", seed=4, params=[6]),
              cg.HostScene.generate("c1", seed=3, params=[0.125, 1, 256])]  # textured
    rng = np.random.default_rng(5)
    reqs = []
    for k in range(12):
        w, h = [(32, 24), (48, 40), (17, 11), (64, 36)][k % 4]
        pos, d = workloads.orbit_camera(rng, 3.0 if k % 2 else 12.0)
        cam = cg.Camera.make(pos, d, (0.5 * w / h, 0.5))
        reqs.append((k % 2, cam, cg.Conf.make(w, h, [1, 2, 4][k % 3], [2, 5, 8][k % 3])))
    with cg.Server(scenes, budget_bytes=256 << 20, engine=engine, max_trace_depth=8, rounds_per_launch=4) as srv:
        ids = {}
        for i, (s, cam, conf) in enumerate(reqs):
            ids[srv.submit(s, cam, conf)] = i
            if i % 3 == 2:
                time.sleep(0.01)  # later requests arrive while earlier ones render
        got = {}
        t0 = time.time()
        while len(got) < len(reqs) and time.time() - t0 < 60:
            r = srv.next(1.0)
            if r is not None:
                got[ids[r["id"]]] = r
        assert len(got) == len(reqs)
        for i, (s, cam, conf) in enumerate(reqs):
            ref = srv.render_alone(s, cam, conf)
            assert got[i]["history"].tobytes() == ref["history"].tobytes(), f"request {i}"
            assert got[i]["latency_s"] >= got[i]["queued_s"] >= 0
        st = srv.stats()
        assert st["submitted"] == len(reqs) and st["completed"] == len(reqs) and st["rays"] > 0


def test_serve_rejects_oversized_request(cg):
    scenes = [cg.HostScene.generate("spheres", seed=4, params=[4])]
    with cg.Server(scenes, budget_bytes=16 << 20, engine=1, max_trace_depth=4) as srv:
        with pytest.raises(cg.NativeError, match="budget"):
            srv.submit(0, scenes[0].describe()["camera"], cg.Conf.make(1920, 1080, 1, 2))
        with pytest.raises(cg.NativeError, match="max_trace_depth"):
            srv.submit(0, scenes[0].describe()["camera"], cg.Conf.make(16, 16, 1, 8))


def test_serve_refuses_beyond_the_waiting_bound(cg):
    """Overload policy: with max_waiting = 3, a burst of 40 requests submitted
    at once is partly refused (submit returns None) and counted; every
    accepted request completes and still equals its alone render."""
    scenes = [cg.HostScene.generate("spheres", seed=4, params=[6])]
    cam = scenes[0].describe()["camera"]
    conf = cg.Conf.make(64, 48, 4, 6)
    with cg.Server(scenes, budget_bytes=256 << 20, engine=0, max_trace_depth=8, in_flight=1, max_waiting=3) as srv:
        accepted = [rid for rid in (srv.submit(0, cam, conf) for _ in range(40)) if rid is not None]
        st = srv.stats()
        assert st["refused"] > 0 and st["submitted"] == 40
        assert len(accepted) + st["refused"] == 40
        got = []
        t0 = time.time()
        while len(got) < len(accepted) and time.time() - t0 < 60:
            r = srv.next(1.0)
            if r is not None:
                got.append(r)
        assert sorted(r["id"] for r in got) == sorted(accepted)
        ref = srv.render_alone(0, cam, conf)
        assert all(r["history"].tobytes() == ref["history"].tobytes() for r in got)
        assert srv.stats()["completed"] == len(accepted)
