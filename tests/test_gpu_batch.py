"""GPU: the batched (wavefront) engine.  Every job of a batch must finish
with exactly the per-pixel result of rendering it alone with cg_view (which
the render parity tests hold to the oracle): history XYZ and carried RNG
state, rgba8, path-record hash and query count, bit for bit -- whether the
rounds run as a captured hipGraph or as plain launches, and whether jobs are
admitted together or while others are in flight."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scenes(cg):
    out = {}
    for key, name, params, seed in (("spheres", "spheres", [8], 3), ("c1", "c1", [0.125], 2), ("c0", "c0", [], 1),
                                    ("spheres_tex", "spheres", [8, 1, 64], 4)):
        s = cg.HostScene.generate(name, seed=seed, params=params)
        out[key] = (s, s.scene_data())
    return out


def _jobs():
    # (scene, width, height, spp, depth, camera perturbation)
    return [("spheres", 40, 24, 3, 4, 0.0), ("c1", 72, 40, 2, 8, 0.3), ("c0", 64, 36, 1, 2, 0.0),
            ("spheres", 17, 9, 2, 3, 0.7), ("c1", 24, 24, 1, 5, -0.4), ("c0", 33, 20, 4, 3, 0.2),
            ("spheres_tex", 48, 32, 2, 4, 0.5)]


def _camera(cg, desc, k):
    c = desc["camera"]
    if k == 0.0:
        return c
    pos = np.array(c.position[:], np.float64)
    ang = k
    rot = np.array([[np.cos(ang), -np.sin(ang), 0], [np.sin(ang), np.cos(ang), 0], [0, 0, 1]])
    p2 = rot @ pos
    d2 = rot @ np.array(c.direction[:], np.float64)
    return cg.Camera.make(p2, d2, c.viewport[:])


def _alone(cg, scenes, job, iters=1):
    name, w, h, spp, depth, k = job
    s, sd = scenes[name]
    conf = cg.Conf.make(w, h, spp, depth)
    cam = _camera(cg, s.describe(), k)
    with cg.DeviceEnv(conf) as env:
        ds = env.upload(sd)
        env.set_path_records(True)
        for it in range(iters):
            env.view(ds, cam, it)
        return env.read_history(), env.read_rgba8(), env.read_path_records()


def _check_equal(got, ref, label):
    hist, rgba, (ph, rc) = got
    rh, rr, (rph, rrc) = ref
    assert hist.tobytes() == rh.tobytes(), f"{label}: history differs"
    assert rgba.tobytes() == rr.tobytes(), f"{label}: rgba differs"
    assert np.array_equal(ph, rph) and np.array_equal(rc, rrc), f"{label}: path records differ"


@pytest.fixture(scope="module")
def scenes(cg):
    return _scenes(cg)


@pytest.fixture(scope="module")
def alone(cg, scenes):
    return [_alone(cg, scenes, j) for j in _jobs()]


def _run_batch(cg, scenes, jobs, use_graph=True, staggered=False, slot_capacity=1 << 16, rounds=8,
               block_size=256, short_stack=16):
    env = cg.DeviceEnv(cg.Conf.make(8, 8))
    ds = {n: env.upload(sd) for n, (s, sd) in scenes.items()}
    b = cg.Batch(env, slot_capacity=slot_capacity, max_trace_depth=8, rounds_per_launch=rounds,
                 use_graph=use_graph, block_size=block_size, short_stack=short_stack)
    rjs = []
    for i, (name, w, h, spp, depth, k) in enumerate(jobs):
        s, _ = scenes[name]
        rj = cg.RenderJob(env, ds[name], _camera(cg, s.describe(), k), cg.Conf.make(w, h, spp, depth),
                          path_records=True)
        assert b.submit(rj)
        rjs.append(rj)
        if staggered and i % 2 == 1:
            b.launch(1)  # later jobs join while earlier ones are in flight
    finished = set()
    for _ in range(10000):
        if len(finished) == len(rjs):
            break
        b.launch(1)
        done, active = b.poll()
        finished.update(done)
    assert len(finished) == len(rjs)
    out = [(rj.read_history(), rj.read_rgba8(), rj.read_path_records()) for rj in rjs]
    st = b.stats()
    for rj in rjs:
        rj.free()
    b.close()
    env.close()
    return out, st


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_equals_view(cg, scenes, alone, use_graph):
    got, st = _run_batch(cg, scenes, _jobs(), use_graph=use_graph)
    for i, (g, r) in enumerate(zip(got, alone)):
        _check_equal(g, r, f"job {i}")
    # every query the path records count was traced by the batch
    total = sum(int(g[2][1].sum()) for g in got)
    assert st["rays"] + st["shadow_rays"] == total


@pytest.mark.parametrize("block_size,short_stack", [(b, k) for b in (64, 128, 256) for k in (8, 16, 24)])
def test_batch_kernel_shapes(cg, scenes, alone, block_size, short_stack):
    """Every kernel shape the autotuner can pick gives the same results (the
    short stack spills to the overflow area at a different depth)."""
    got, _ = _run_batch(cg, scenes, _jobs()[:3], block_size=block_size, short_stack=short_stack)
    for i, (g, r) in enumerate(zip(got, alone[:3])):
        _check_equal(g, r, f"job {i}")


def test_batch_staggered_admission(cg, scenes, alone):
    got, _ = _run_batch(cg, scenes, _jobs(), staggered=True)
    for i, (g, r) in enumerate(zip(got, alone)):
        _check_equal(g, r, f"job {i}")


def test_batch_waits_for_slots(cg, scenes):
    """A batch with room for one job at a time refuses the second until the
    first has finished, then renders it to the same result."""
    env = cg.DeviceEnv(cg.Conf.make(8, 8))
    s, sd = scenes["spheres"]
    ds = env.upload(sd)
    cam = s.describe()["camera"]
    conf = cg.Conf.make(32, 32, 2, 3)
    b = cg.Batch(env, slot_capacity=1024, max_trace_depth=4, rounds_per_launch=4)
    j1 = cg.RenderJob(env, ds, cam, conf, path_records=True)
    j2 = cg.RenderJob(env, ds, cam, conf, path_records=True)
    assert b.submit(j1)
    assert not b.submit(j2)
    b.drain()
    assert b.submit(j2)
    b.drain()
    h1, h2 = j1.read_history(), j2.read_history()
    assert h1.tobytes() == h2.tobytes()
    for j in (j1, j2):
        j.free()
    b.close()
    env.close()


def test_batch_progressive_history(cg, scenes):
    """A job with iter_count > 0 continues an accumulation from its history,
    as cg_view does on its second frame."""
    job = ("spheres", 24, 16, 2, 3, 0.0)
    ref = _alone(cg, scenes, job, iters=2)
    name, w, h, spp, depth, _ = job
    s, sd = scenes[name]
    env = cg.DeviceEnv(cg.Conf.make(8, 8))
    ds = env.upload(sd)
    b = cg.Batch(env, slot_capacity=4096, max_trace_depth=4, rounds_per_launch=4)
    conf = cg.Conf.make(w, h, spp, depth)
    j0 = cg.RenderJob(env, ds, s.describe()["camera"], conf, path_records=True)
    assert b.submit(j0)
    b.drain()
    j1 = cg.RenderJob(env, ds, s.describe()["camera"], conf, iter_count=1, path_records=True,
                      history=j0.read_history())
    assert b.submit(j1)
    b.drain()
    _check_equal((j1.read_history(), j1.read_rgba8(), j1.read_path_records()), ref, "second frame")
    for j in (j0, j1):
        j.free()
    b.close()
    env.close()


def test_view_batch_equals_view(cg, scenes, alone):
    """All jobs in one persistent launch of the per-pixel kernel (and a second
    launch overlapping the first on another stream) give cg_view's results."""
    env = cg.DeviceEnv(cg.Conf.make(8, 8))
    ds = {n: env.upload(sd) for n, (s, sd) in scenes.items()}
    mk = lambda: [cg.RenderJob(env, ds[name], _camera(cg, scenes[name][0].describe(), k), cg.Conf.make(w, h, spp, d),
                               path_records=True) for (name, w, h, spp, d, k) in _jobs()]
    a, b = mk(), mk()
    va = cg.ViewBatch(env, a)
    vb = cg.ViewBatch(env, b[::-1])  # another order, in flight at the same time
    va.wait()
    vb.wait()
    assert va.done() and vb.done()
    va.release()
    vb.release()
    for jobs in (a, b):
        for i, (rj, r) in enumerate(zip(jobs, alone)):
            _check_equal((rj.read_history(), rj.read_rgba8(), rj.read_path_records()), r, f"job {i}")
    for rj in a + b:
        rj.free()
    env.close()
