"""GPU: cg_trace_closest_noalloc, the closest-hit batch that allocates no
device memory, against cg_trace_closest through the C-ABI.

Bar: bit-exact (hit t, instance, triangle and barycentric uv) on two calls in a
row, the second with a smaller input, on scenes whose traversal stacks spill
past the LDS part into the fixed scratch; nothing is written past n; the
device's free memory does not change across the calls."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _free_bytes() -> int:
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


@pytest.mark.parametrize("name,params,accel", [("c1", [0.25], 0), ("c0", [], 1), ("spheres", [8], 2)])
def test_noalloc_matches_trace_closest(cg, name, params, accel):
    env = cg.DeviceEnv(cg.Conf.make(64, 64, 1, 1))
    try:
        warm = cg.HostScene.generate("quad", seed=1)
        wds = env.upload(warm.scene_data())
        env.trace_closest(wds, warm.rays(256, seed=1))  # loads the code objects
        s = cg.HostScene.generate(name, seed=11, params=params)
        env.set_accel(accel)
        ds = env.upload(s.scene_data())
        env.set_accel(1)
        _, depth = ds.footprint()
        print(f"{name} accel {accel}: traversal stack {depth} entries", flush=True)
        assert depth > 16  # the overflow part of the stack is in the scratch
        sizes = [1 << 16, 3000]
        rays = [s.rays(n, seed=3 + i) for i, n in enumerate(sizes)]
        bufs = []
        for r in rays:
            n = r.shape[0]
            rb = env.alloc(n * cg.RAY.itemsize).upload(r)
            hb = env.alloc((n + 64) * cg.HIT.itemsize)
            ub = env.alloc((n + 64) * 8)
            hb.upload(np.full((n + 64) * cg.HIT.itemsize, 0xAB, np.uint8))
            ub.upload(np.full((n + 64) * 8, 0xCD, np.uint8))
            bufs.append((n, rb, hb, ub))
        env.sync()
        free0 = _free_bytes()
        for n, rb, hb, ub in bufs:
            env.trace_closest_noalloc_device(ds, rb, n, hb, ub)
        env.sync()
        free1 = _free_bytes()
        assert abs(free0 - free1) < (8 << 20), f"free device memory moved by {free0 - free1} bytes"
        for (n, rb, hb, ub), r in zip(bufs, rays):
            got = hb.download(np.uint8, (n + 64) * cg.HIT.itemsize)
            got_uv = ub.download(np.uint8, (n + 64) * 8)
            assert np.all(got[n * cg.HIT.itemsize:] == 0xAB) and np.all(got_uv[n * 8:] == 0xCD)
            ref, ref_uv = env.trace_closest(ds, r, with_uv=True)
            hits = got[: n * cg.HIT.itemsize].view(cg.HIT)
            uv = got_uv[: n * 8].view(np.float32).reshape(n, 2)
            hit_frac = float((ref["inst"] != 0xFFFFFFFF).mean())
            print(f"  n={n}: hits {hit_frac * 100:.1f}%", flush=True)
            assert hit_frac > 0.05
            assert hits.tobytes() == ref.tobytes()
            assert uv.tobytes() == ref_uv.tobytes()
        for _, rb, hb, ub in bufs:
            rb.free()
            hb.free()
            ub.free()
        ds.release()
        wds.release()
    finally:
        env.close()


def test_noalloc_empty_and_null(cg):
    env = cg.DeviceEnv(cg.Conf.make(8, 8, 1, 1))
    try:
        s = cg.HostScene.generate("spheres", seed=2, params=[4])
        ds = env.upload(s.scene_data())
        from crystalgraphics_amd import native
        d = native.device()
        assert d.cg_trace_closest_noalloc(env.handle, ds.handle, None, 0, None, None) == 0
        assert d.cg_trace_closest_noalloc(env.handle, ds.handle, None, 5, None, None) == 1
        ds.release()
    finally:
        env.close()
