"""crystalgraphics_amd — MI355X (gfx950) spectral path-tracing core.

Python face of the native libraries:
  * the crystal::graphics API (EnvInit, EnvTerm, EnvPoll, EnvPresent,
    EnvStatus, LoadScene, View, ViewToBuffer), as the reference's
    include/CrystalGraphics headers declare it;
  * HostScene / SceneData: procedural scenes and their flattened SceneData;
  * DeviceEnv / DeviceScene: the C-ABI of include/cg_device.h (upload, the
    per-pixel compute pass, batch closest-hit and occlusion queries).
Every compute call runs on the GPU through libcg_device.so; there is no
fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import native
from .native import (ALIAS, BLAS_NODE, EMISSIVE, HISTORY, HIT, INDEX, INSTANCE, MATERIAL, MISS_ID, MISS_T,
                     PRIMITIVE, RAY, TLAS_NODE, VERTEX, Camera, Conf, NativeError)

__all__ = ["HostScene", "SceneData", "DeviceEnv", "DeviceScene", "Camera", "Conf", "NativeError", "Batch", "RenderJob", "Server", "ViewBatch",
           "EnvInit", "EnvTerm", "EnvPoll", "EnvPresent", "EnvStatus", "LoadScene", "View", "ViewToBuffer",
           "MISS_ID", "MISS_T", "RAY", "HIT", "HISTORY", "VERTEX", "MATERIAL", "PRIMITIVE", "TLAS_NODE",
           "INSTANCE", "BLAS_NODE", "INDEX", "EMISSIVE", "ALIAS"]


def _host_check(rc: int) -> None:
    if rc != 0:
        raise NativeError(rc, native.host().cgh_last_error().decode())


class SceneData:
    """A flattened scene (the reference's SceneData) owned by the host library."""

    def __init__(self, handle: int, scene: "HostScene"):
        self._h = handle
        self._scene = scene  # vertices and materials are referenced, keep alive
        self._c = native.SceneDataC()
        native.host().cgh_scene_data_view(self._h, C.byref(self._c))

    @property
    def c(self) -> native.SceneDataC:
        return self._c

    def arrays(self) -> dict:
        return native.scene_data_arrays(self._c)

    def close(self) -> None:
        if self._h:
            native.host().cgh_scene_data_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HostScene:
    """A scene held by the host layer (the reference's Scene)."""

    def __init__(self, handle: int):
        self._h = handle

    @classmethod
    def generate(cls, name: str, seed: int = 1, params=()) -> "HostScene":
        h = C.c_void_p()
        p = np.asarray(params, dtype=np.float64)
        rc = native.host().cgh_scene_generate(name.encode(), seed, p.ctypes.data if p.size else None,
                                              int(p.size), C.byref(h))
        _host_check(rc)
        return cls(h.value)

    @classmethod
    def from_arrays(cls, vertices, indices, materials, prims, env=None, tag: int = 1, textures=None) -> "HostScene":
        """textures: list of {"rgba8": (h, w, 4) uint8, "wrap_s", "wrap_t"}, indexed by the materials'
        texture slots."""
        v = np.ascontiguousarray(vertices, VERTEX)
        i = np.ascontiguousarray(indices, np.uint32)
        m = np.ascontiguousarray(materials, MATERIAL)
        p = np.ascontiguousarray(prims, PRIMITIVE)
        e = None if env is None else np.ascontiguousarray(env, np.float32)
        tex, keep = native.texture_structs(textures)
        h = C.c_void_p()
        rc = native.host().cgh_scene_create(
            v.ctypes.data, v.shape[0], i.ctypes.data, i.size, m.ctypes.data, m.shape[0], p.ctypes.data,
            p.shape[0], tex, len(keep), None if e is None else e.ctypes.data, 0 if e is None else e.shape[1],
            0 if e is None else e.shape[0], tag, C.byref(h))
        del keep
        _host_check(rc)
        return cls(h.value)

    def describe(self) -> dict:
        d = native.SceneDescC()
        native.host().cgh_scene_describe(self._h, C.byref(d))
        out = {
            "vertices": native.view_array(d.vertices, d.vertex_count, VERTEX),
            "indices": native.view_array(d.indices, d.index_count, np.dtype("<u4")),
            "materials": native.view_array(d.materials, d.material_count, MATERIAL),
            "prims": native.view_array(d.prims, d.prim_count, PRIMITIVE),
            "env": None if not d.env else native.view_array(d.env, d.env_w * d.env_h * 4, np.dtype("<f4")).reshape(
                d.env_h, d.env_w, 4),
            "camera": d.camera, "conf": d.conf, "tag": d.tag,
            "textures": [],
        }
        if d.texture_count:
            views = (native.TextureC * d.texture_count).from_address(d.textures)
            for t in views:
                img = native.view_array(t.rgba8, t.width * t.height * 4, np.dtype("u1")).reshape(t.height, t.width, 4)
                out["textures"].append({"rgba8": img, "wrap_s": t.wrap_s, "wrap_t": t.wrap_t})
        return out

    def scene_data(self, share_blas: bool = True, threads: int = 8) -> SceneData:
        h = C.c_void_p()
        _host_check(native.host().cgh_scene_data_build(self._h, int(share_blas), int(threads), C.byref(h)))
        return SceneData(h.value, self)

    def rays(self, n: int, seed: int = 1) -> np.ndarray:
        out = np.zeros(n, RAY)
        _host_check(native.host().cgh_rays_generate(self._h, seed, n, out.ctypes.data))
        return out

    @property
    def handle(self) -> int:
        return self._h

    def close(self) -> None:
        if self._h:
            native.host().cgh_scene_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceScene:
    def __init__(self, env: "DeviceEnv", handle: int):
        self.env = env
        self._h = handle

    @property
    def handle(self) -> int:
        return self._h

    def footprint(self) -> tuple[int, int]:
        b = C.c_uint64()
        s = C.c_uint32()
        self.env._check(native.device().cg_scene_footprint(self._h, C.byref(b), C.byref(s)))
        return int(b.value), int(s.value)

    def release(self) -> None:
        if self._h and self.env._h:
            self.env._check(native.device().cg_scene_release(self.env._h, self._h))
        self._h = None


class DeviceBuffer:
    """An HBM allocation made through cg_device_alloc."""

    def __init__(self, env: "DeviceEnv", nbytes: int):
        self.env = env
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        env._check(native.device().cg_device_alloc(env._h, self.nbytes, C.byref(p)))
        self.ptr = p.value

    def upload(self, arr: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        self.env._check(native.device().cg_memcpy_h2d(self.env._h, self.ptr, a.ctypes.data, a.nbytes))
        return self

    def download(self, dtype, count: int) -> np.ndarray:
        out = np.zeros(count, dtype)
        assert out.nbytes <= self.nbytes
        self.env._check(native.device().cg_memcpy_d2h(self.env._h, out.ctypes.data, self.ptr, out.nbytes))
        return out

    def free(self) -> None:
        if self.ptr:
            self.env._check(native.device().cg_device_free(self.env._h, self.ptr))
            self.ptr = None


class DeviceEnv:
    """The HIP backend (cg_env): one stream, a history/target, scenes."""

    def __init__(self, conf: Conf | None = None, device: int = 0):
        self.conf = conf or Conf.make()
        h = C.c_void_p()
        rc = native.device().cg_env_create(C.byref(self.conf), device, C.byref(h))
        if rc != 0:
            raise NativeError(rc, native.device().cg_last_error().decode())
        self._h = h.value

    def _check(self, rc: int) -> None:
        if rc != 0:
            msg = native.device().cg_env_last_error(self._h).decode() if self._h else \
                native.device().cg_last_error().decode()
            raise NativeError(rc, msg)

    @property
    def handle(self) -> int:
        return self._h

    @property
    def stream(self) -> int:
        return native.device().cg_env_stream(self._h)

    def upload(self, sd) -> DeviceScene:
        """Upload a SceneData (or a dict of arrays in SceneData layout)."""
        keep = None
        if isinstance(sd, SceneData):
            c = sd.c
        else:
            c, keep = native.make_scene_data(sd)
        h = C.c_void_p()
        self._check(native.device().cg_scene_upload(self._h, C.byref(c), C.byref(h)))
        del keep
        return DeviceScene(self, h.value)

    ACCEL_REFERENCE, ACCEL_REBUILT = 0, 1

    def refit(self, scene: DeviceScene, sd) -> None:
        """New data (same triangles and BLASes) for a scene uploaded with accel 2: BLAS boxes refitted on the
        device, TLAS and other records rebuilt."""
        keep = None
        if isinstance(sd, SceneData):
            c = sd.c
        else:
            c, keep = native.make_scene_data(sd)
        self._check(native.device().cg_scene_refit(self._h, scene.handle, C.byref(c)))
        del keep

    def set_accel(self, mode: int) -> None:
        """Acceleration layout for later uploads (0: the reference's trees, 1: rebuilt SAH, 2: every BLAS
        built on the device)."""
        self._check(native.device().cg_env_set_accel(self._h, int(mode)))

    RNG_REFERENCE, RNG_SAMPLE = 0, 1

    def set_rng_mode(self, mode: int) -> None:
        """RNG_REFERENCE: the reference's one stream per pixel; RNG_SAMPLE: one stream per sample (cg_device.h)."""
        self._check(native.device().cg_env_set_rng_mode(self._h, mode))

    def set_schedule(self, on: bool = True) -> None:
        """Work order of view/view_job: longest first by the last same-size render's durations (on), or image order."""
        self._check(native.device().cg_env_set_schedule(self._h, int(on)))

    def set_counting(self, on: bool = True) -> None:
        self._check(native.device().cg_env_set_counting(self._h, int(on)))

    def set_path_records(self, on: bool = True) -> None:
        self._check(native.device().cg_env_set_path_records(self._h, int(on)))

    def stats(self) -> dict:
        s = native.Stats()
        self._check(native.device().cg_stats_read(self._h, C.byref(s)))
        return {k: int(getattr(s, k)) for k, _ in native.Stats._fields_}

    def reset_stats(self) -> None:
        self._check(native.device().cg_stats_reset(self._h))

    def sync(self) -> None:
        self._check(native.device().cg_env_sync(self._h))

    def view(self, scene: DeviceScene, camera: Camera, iter_count: int = 0) -> None:
        self._check(native.device().cg_view(self._h, scene.handle, C.byref(camera), iter_count))

    def read_history(self) -> np.ndarray:
        w, h = self.conf.render_width, self.conf.render_height
        out = np.zeros(w * h, HISTORY)
        self._check(native.device().cg_read_target(self._h, out.ctypes.data, None))
        return out.reshape(h, w)

    def read_rgba8(self) -> np.ndarray:
        w, h = self.conf.render_width, self.conf.render_height
        out = np.zeros((h, w, 4), np.uint8)
        self._check(native.device().cg_read_target(self._h, None, out.ctypes.data))
        return out

    def read_path_records(self) -> tuple[np.ndarray, np.ndarray]:
        w, h = self.conf.render_width, self.conf.render_height
        hs = np.zeros(w * h, np.uint32)
        rc = np.zeros(w * h, np.uint32)
        self._check(native.device().cg_read_path_records(self._h, hs.ctypes.data, rc.ctypes.data))
        return hs.reshape(h, w), rc.reshape(h, w)

    def write_history(self, hist: np.ndarray) -> None:
        a = np.ascontiguousarray(hist, HISTORY)
        self._check(native.device().cg_write_history(self._h, a.ctypes.data))

    def stream_copy(self, dst: DeviceBuffer, src: DeviceBuffer, nbytes: int) -> None:
        self._check(native.device().cg_stream_copy(self._h, dst.ptr, src.ptr, nbytes))

    def alias_pick(self, scene: DeviceScene, u: np.ndarray) -> np.ndarray:
        """The emitters the device's alias sampling picks for uniform pairs u (n x 2 float32)."""
        u = np.ascontiguousarray(u, np.float32).reshape(-1, 2)
        n = u.shape[0]
        ub = self.alloc(max(n, 1) * 8).upload(u)
        ob = self.alloc(max(n, 1) * 4)
        try:
            self._check(native.device().cg_alias_pick(self._h, scene.handle, ub.ptr, n, ob.ptr))
            return ob.download(np.uint32, n)
        finally:
            ub.free()
            ob.free()

    def stream_read(self, src: DeviceBuffer, nbytes: int, sink: DeviceBuffer) -> None:
        self._check(native.device().cg_stream_read(self._h, src.ptr, nbytes, sink.ptr))

    def alloc(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def trace_closest_device(self, scene: DeviceScene, rays: DeviceBuffer, n: int, hits: DeviceBuffer,
                             uv: DeviceBuffer | None = None) -> None:
        self._check(native.device().cg_trace_closest(self._h, scene.handle, rays.ptr, n, hits.ptr,
                                                     uv.ptr if uv else None))

    def trace_closest_noalloc_device(self, scene: DeviceScene, rays: DeviceBuffer, n: int, hits: DeviceBuffer,
                                     uv: DeviceBuffer | None = None) -> None:
        """cg_trace_closest_noalloc: the closest-hit batch with no device allocation inside the call."""
        self._check(native.device().cg_trace_closest_noalloc(self._h, scene.handle, rays.ptr, n, hits.ptr,
                                                             uv.ptr if uv else None))

    def trace_closest(self, scene: DeviceScene, rays: np.ndarray, with_uv: bool = False):
        """Host convenience: copy rays in, closest hits (and uv) out."""
        n = rays.shape[0]
        rb = self.alloc(max(n, 1) * RAY.itemsize).upload(rays)
        hb = self.alloc(max(n, 1) * HIT.itemsize)
        ub = self.alloc(max(n, 1) * 8) if with_uv else None
        try:
            self.trace_closest_device(scene, rb, n, hb, ub)
            hits = hb.download(HIT, n)
            uv = ub.download(np.float32, 2 * n).reshape(n, 2) if ub else None
        finally:
            rb.free()
            hb.free()
            if ub:
                ub.free()
        return (hits, uv) if with_uv else hits

    def trace_occluded(self, scene: DeviceScene, rays: np.ndarray, tmax: np.ndarray) -> np.ndarray:
        n = rays.shape[0]
        rb = self.alloc(max(n, 1) * RAY.itemsize).upload(rays)
        tb = self.alloc(max(n, 1) * 4).upload(np.asarray(tmax, np.float32))
        ob = self.alloc(max(n, 1) * 4)
        try:
            self._check(native.device().cg_trace_occluded(self._h, scene.handle, rb.ptr, tb.ptr, n, ob.ptr))
            return ob.download(np.uint32, n)
        finally:
            rb.free()
            tb.free()
            ob.free()

    def close(self) -> None:
        if self._h:
            native.device().cg_env_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RenderJob:
    """One render job of a Batch: a progressive frame of one scene and camera
    at its own conf, with device buffers for its history (in/out), rgba8
    target and path records."""

    def __init__(self, env: DeviceEnv, scene: DeviceScene, camera: Camera, conf: Conf, iter_count: int = 0,
                 path_records: bool = False, history: np.ndarray | None = None, rng_mode: int = 0):
        self.env, self.scene, self.camera, self.conf, self.iter_count = env, scene, camera, conf, iter_count
        self.rng_mode = rng_mode
        n = conf.render_width * conf.render_height
        self.npix = n
        self.history = env.alloc(n * HISTORY.itemsize)
        if history is not None:
            self.history.upload(np.ascontiguousarray(history, HISTORY).reshape(-1))
        self.rgba = env.alloc(n * 4)
        self.path_hash = env.alloc(n * 4) if path_records else None
        self.ray_count = env.alloc(n * 4) if path_records else None
        self.id = None

    def c(self) -> native.Job:
        j = native.Job()
        j.scene = self.scene.handle
        j.camera = self.camera
        j.conf = self.conf
        j.iter_count = self.iter_count
        j.rng_mode = self.rng_mode
        j.history = self.history.ptr
        j.rgba8 = self.rgba.ptr
        j.path_hash = self.path_hash.ptr if self.path_hash else None
        j.ray_count = self.ray_count.ptr if self.ray_count else None
        return j

    def view_alone(self) -> None:
        """Render this job with the plain one-lane-per-pixel kernel (cg_view_job)."""
        self.env._check(native.device().cg_view_job(self.env.handle, C.byref(self.c())))

    def read_history(self) -> np.ndarray:
        return self.history.download(HISTORY, self.npix).reshape(self.conf.render_height, self.conf.render_width)

    def read_rgba8(self) -> np.ndarray:
        return self.rgba.download(np.uint8, self.npix * 4).reshape(self.conf.render_height, self.conf.render_width, 4)

    def read_path_records(self) -> tuple[np.ndarray, np.ndarray]:
        h, w = self.conf.render_height, self.conf.render_width
        return (self.path_hash.download(np.uint32, self.npix).reshape(h, w),
                self.ray_count.download(np.uint32, self.npix).reshape(h, w))

    def free(self) -> None:
        for b in (self.history, self.rgba, self.path_hash, self.ray_count):
            if b:
                b.free()


class ViewBatch:
    """Many jobs in one persistent launch of the plain per-pixel kernel
    (cg_view_batch_launch), on a stream of its own; results equal
    RenderJob.view_alone's."""

    def __init__(self, env: DeviceEnv, jobs: list[RenderJob]):
        self.env, self.jobs = env, list(jobs)
        arr = (native.Job * len(self.jobs))(*[j.c() for j in self.jobs])
        h = C.c_void_p()
        env._check(native.device().cg_view_batch_launch(env.handle, arr, len(self.jobs), C.byref(h)))
        self._h = h.value

    def done(self) -> bool:
        d = C.c_int32()
        self.env._check(native.device().cg_view_batch_query(self._h, C.byref(d)))
        return bool(d.value)

    def wait(self) -> None:
        self.env._check(native.device().cg_view_batch_wait(self._h))

    def release(self) -> None:
        if self._h and self.env._h:  # a closed env has released its view batches
            self.env._check(native.device().cg_view_batch_release(self._h))
        self._h = None


class Batch:
    """The wavefront engine for many render jobs (cg_batch): jobs are admitted
    between launches of `rounds_per_launch` rounds; each finishes with the
    same per-pixel result as DeviceEnv.view."""

    def __init__(self, env: DeviceEnv, slot_capacity: int = 1 << 20, max_trace_depth: int = 8,
                 rounds_per_launch: int = 16, use_graph: bool = True, max_jobs: int = 256,
                 max_stack_depth: int = 128, block_size: int = 256, short_stack: int = 16):
        self.env = env
        self.conf = native.BatchConf(slot_capacity, max_trace_depth, rounds_per_launch, int(use_graph), max_jobs,
                                     max_stack_depth, block_size, short_stack)
        h = C.c_void_p()
        env._check(native.device().cg_batch_create(env.handle, C.byref(self.conf), C.byref(h)))
        self._h = h.value

    def submit(self, job: RenderJob) -> bool:
        """Admit a job; False when there is no room now (CG_ERR_FULL)."""
        jid = C.c_uint32()
        rc = native.device().cg_batch_submit(self._h, C.byref(job.c()), C.byref(jid))
        if rc == 6:
            return False
        self.env._check(rc)
        job.id = jid.value
        return True

    def launch(self, launches: int = 1) -> None:
        self.env._check(native.device().cg_batch_launch(self._h, launches))

    def poll(self, cap: int = 4096) -> tuple[list[int], int]:
        done = (C.c_uint32 * cap)()
        nd, na = C.c_uint32(), C.c_uint32()
        self.env._check(native.device().cg_batch_poll(self._h, done, cap, C.byref(nd), C.byref(na)))
        return [done[i] for i in range(nd.value)], na.value

    def drain(self) -> None:
        self.env._check(native.device().cg_batch_drain(self._h))

    def stats(self, reset: bool = False) -> dict:
        s = native.Stats()
        self.env._check(native.device().cg_batch_stats(self._h, C.byref(s), int(reset)))
        return {k: int(getattr(s, k)) for k, _ in native.Stats._fields_}

    def footprint(self) -> tuple[int, int]:
        b, f = C.c_uint64(), C.c_uint32()
        self.env._check(native.device().cg_batch_footprint(self._h, C.byref(b), C.byref(f)))
        return b.value, f.value

    def close(self) -> None:
        if self._h and self.env._h:  # a closed env has destroyed its batches
            native.device().cg_batch_destroy(self._h)
        self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Server:
    """The serving form (crystal::graphics::Server, include/CrystalGraphics/
    serve.h): scenes resident in HBM, requests admitted within an HBM budget
    and batched as they arrive; each result equals a plain View's."""

    MAX_PIXELS = 1920 * 1080

    LAUNCHES, WAVEFRONT = 0, 1

    def __init__(self, scenes: list[HostScene], budget_bytes: int = 16 << 30, engine: int = 0,
                 output_fraction: float = 0.25, max_trace_depth: int = 8, in_flight: int = 4,
                 rounds_per_launch: int = 16, block_size: int = 0, short_stack: int = 0, max_waiting: int = 0,
                 rng_mode: int = 0):
        arr = (C.c_void_p * len(scenes))(*[s.handle for s in scenes])
        h = C.c_void_p()
        rc = native.host().cgh_server_create(arr, len(scenes), engine, int(budget_bytes), float(output_fraction),
                                             max_trace_depth, in_flight, rounds_per_launch, block_size, short_stack,
                                             max_waiting, rng_mode, C.byref(h))
        if rc != 0:
            raise NativeError(rc, native.host().cgh_serve_last_error().decode())
        self._h = h.value
        self._buf = np.zeros(self.MAX_PIXELS, HISTORY)

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise NativeError(rc, native.host().cgh_serve_last_error().decode())

    def submit(self, scene: int, camera: Camera, conf: Conf) -> int | None:
        """The request's id, or None when the server refuses it (max_waiting requests already wait)."""
        rid = C.c_uint64()
        rc = native.host().cgh_server_submit(self._h, scene, C.byref(camera), C.byref(conf), C.byref(rid))
        if rc == 6:
            return None
        self._check(rc)
        return rid.value

    def _result(self, info) -> dict:
        n = info.width * info.height
        return {"id": info.id, "history": self._buf[:n].copy().reshape(info.height, info.width),
                "queued_s": info.queued_s, "latency_s": info.latency_s}

    def next(self, timeout: float = 1.0) -> dict | None:
        """The next finished request, or None if none finished in time."""
        info = native.ResultInfo()
        self._check(native.host().cgh_server_next(self._h, timeout, C.byref(info), self._buf.ctypes.data,
                                                  self._buf.shape[0]))
        return self._result(info) if info.id else None

    def render_alone(self, scene: int, camera: Camera, conf: Conf) -> dict:
        info = native.ResultInfo()
        self._check(native.host().cgh_server_render_alone(self._h, scene, C.byref(camera), C.byref(conf),
                                                          C.byref(info), self._buf.ctypes.data, self._buf.shape[0]))
        return self._result(info)

    def stats(self) -> dict:
        s = native.ServeStats()
        self._check(native.host().cgh_server_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in native.ServeStats._fields_}

    def close(self) -> None:
        if self._h:
            native.host().cgh_server_free(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# --- the crystal::graphics API (include/CrystalGraphics/*.h) ----------------
def EnvInit(conf: Conf | None = None) -> None:
    _host_check(native.host().cgh_env_init(C.byref(conf or Conf.make())))


def EnvTerm() -> None:
    _host_check(native.host().cgh_env_term())


def EnvPoll() -> None:
    _host_check(native.host().cgh_env_poll())


def EnvPresent() -> None:
    if not EnvStatus():
        raise NativeError(5, "Environment not initialized.")


def EnvStatus() -> bool:
    return bool(native.host().cgh_env_status())


def LoadScene(path: str) -> HostScene:
    h = C.c_void_p()
    _host_check(native.host().cgh_load_scene(str(path).encode(), C.byref(h)))
    return HostScene(h.value)


def SubmitJobs(scenes: list, requests: list) -> tuple[list[np.ndarray], float]:
    """Batch form through the EnvInit environment: requests are (scene index, Camera, Conf); every request
    rendered in one launch (a new accumulation each).  Returns the (H, W, 3) XYZ images in request order and
    the batch's wall time in seconds."""
    n = len(requests)
    sc = (C.c_void_p * max(1, len(scenes)))(*[s.handle for s in scenes])
    idx = np.array([r[0] for r in requests], np.uint32)
    cams = (Camera * max(1, n))(*[r[1] for r in requests])
    confs = (Conf * max(1, n))(*[r[2] for r in requests])
    outs = [np.zeros((r[2].render_height, r[2].render_width, 3), np.float32) for r in requests]
    ptrs = (C.c_void_p * max(1, n))(*[o.ctypes.data for o in outs])
    secs = C.c_double()
    _host_check(native.host().cgh_submit_jobs(sc, len(scenes), idx.ctypes.data, cams, confs, n, ptrs, C.byref(secs)))
    return outs, secs.value


def View(scene: HostScene, camera: Camera | None = None) -> None:
    _host_check(native.host().cgh_view(scene.handle, C.byref(camera or Camera.make((0, 0, 0), (1, 0, 0)))))


def ViewToBuffer(scene: HostScene, camera: Camera, out: np.ndarray) -> np.ndarray:
    assert out.dtype == np.float32 and out.flags.c_contiguous
    _host_check(native.host().cgh_view_to_buffer(scene.handle, C.byref(camera), out.ctypes.data, out.size))
    return out
