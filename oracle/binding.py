"""ORACLE binding — test infrastructure only.

ctypes access to liboracle.so (the CPU restatement of the reference's path,
cg_oracle.cpp).  Imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from crystalgraphics_amd import native
from crystalgraphics_amd.native import (HISTORY, HIT, MATERIAL, PRIMITIVE, RAY, VERTEX, Camera, Conf,
                                        SceneDataC)

_PATH = Path(__file__).resolve().parent / "liboracle.so"
_lib = None
P = C.c_void_p
F = C.c_float
U32 = C.c_uint32
FP = C.POINTER(C.c_float)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not _PATH.exists():
            raise ImportError(f"{_PATH} is not built (make -C oracle)")
        L = C.CDLL(str(_PATH))
        sig = [
            ("orc_scene_build", P, [P, U32, P, U32, P, U32, P, U32, P, U32, P, U32, U32]),
            ("orc_scene_wrap", P, [C.POINTER(SceneDataC)]),
            ("orc_scene_free", None, [P]),
            ("orc_scene_get", None, [P, C.POINTER(SceneDataC)]),
            ("orc_trace_closest", None, [P, P, U32, P, P, P, P, C.c_int, C.c_int]),
            ("orc_render", C.c_int, [P, C.POINTER(Conf), C.POINTER(Camera), U32, P, P, P, P, U32, U32, C.c_int,
                                     C.c_int]),
            ("orc_sincos", None, [F, FP, FP]),
            ("orc_texture_sample", None, [P, U32, U32, U32, C.c_int32, F, F, FP]),
            ("orc_emissive_texel", None, [C.POINTER(native.TextureC), F, F, FP]),
            ("orc_rng_step", U32, [U32]),
            ("orc_rand_float", F, [C.POINTER(U32)]),
            ("orc_erf", F, [F]),
            ("orc_xyz_match", None, [F, FP]),
            ("orc_xyz_match_sum_integral", F, [F, F]),
            ("orc_sample_wavelength", None, [F, FP, FP]),
            ("orc_illuminate", None, [F, F, FP]),
            ("orc_smits", F, [C.c_int, F]),
            ("orc_rgb_to_reflectance", F, [FP, F]),
            ("orc_rgb_to_radiance", F, [FP, F]),
            ("orc_dispersed_ior", F, [F, F, F]),
            ("orc_fresnel", F, [FP, FP, F, F]),
            ("orc_alias_build", None, [P, U32, P]),
            ("orc_alias_pick", U32, [P, U32, F, F]),
            ("orc_srgb", None, [FP, FP]),
            ("orc_camera_ray", None, [C.POINTER(Camera), U32, U32, U32, U32, P]),
            ("orc_intersect_triangle", C.c_int, [P, FP, FP, FP, FP]),
            ("orc_intersect_aabb", C.c_int, [P, FP, FP]),
            ("orc_sample_lobe", F, [C.c_int, FP, FP, F, F, F, F, C.POINTER(U32), FP]),
        ]
        for name, res, args in sig:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def texture_sample(layers: np.ndarray, layer: int, u: float, v: float) -> np.ndarray:
    """The oracle's material texture lookup; layers is (L, H, W, 4) uint8 (L may be 0)."""
    a = np.ascontiguousarray(layers, np.uint8)
    out = (C.c_float * 4)()
    lib().orc_texture_sample(a.ctypes.data if a.size else None, a.shape[2], a.shape[1], a.shape[0], layer, u, v, out)
    return np.array(out[:], np.float32)


def emissive_texel(image: np.ndarray, wrap_s: int, wrap_t: int, u: float, v: float) -> np.ndarray:
    """The emissive buffer's nearest lookup in a (H, W, 4) uint8 image."""
    img = np.ascontiguousarray(image, np.uint8)
    t = native.TextureC(img.ctypes.data, img.shape[1], img.shape[0], wrap_s, wrap_t)
    out = (C.c_float * 3)()
    lib().orc_emissive_texel(C.byref(t), u, v, out)
    return np.array(out[:], np.float32)


def fvec(values) -> C.Array:
    a = (C.c_float * len(values))(*[float(v) for v in values])
    return a


class OracleScene:
    """A reference-format SceneData held by the oracle."""

    def __init__(self, handle: int):
        self._h = handle

    @classmethod
    def build(cls, desc: dict) -> "OracleScene":
        """Run the oracle's restatement of the reference's scene preparation."""
        v = np.ascontiguousarray(desc["vertices"], VERTEX)
        i = np.ascontiguousarray(desc["indices"], np.uint32)
        m = np.ascontiguousarray(desc["materials"], MATERIAL)
        p = np.ascontiguousarray(desc["prims"], PRIMITIVE)
        e = desc.get("env")
        e = None if e is None else np.ascontiguousarray(e, np.float32)
        tex, keep = native.texture_structs(desc.get("textures"))
        h = lib().orc_scene_build(v.ctypes.data, v.shape[0], i.ctypes.data, i.size, m.ctypes.data, m.shape[0],
                                  p.ctypes.data, p.shape[0], tex, len(keep), None if e is None else e.ctypes.data,
                                  0 if e is None else e.shape[1], 0 if e is None else e.shape[0])
        del keep
        return cls(h)

    @classmethod
    def wrap(cls, sd: SceneDataC) -> "OracleScene":
        return cls(lib().orc_scene_wrap(C.byref(sd)))

    @classmethod
    def wrap_arrays(cls, arrays: dict) -> "OracleScene":
        sd, keep = native.make_scene_data(arrays)
        s = cls.wrap(sd)
        del keep
        return s

    def arrays(self) -> dict:
        sd = SceneDataC()
        lib().orc_scene_get(self._h, C.byref(sd))
        return native.scene_data_arrays(sd)

    def trace_closest(self, rays: np.ndarray, threads: int = 8, cull: bool = True, counts: bool = False):
        n = rays.shape[0]
        r = np.ascontiguousarray(rays, RAY)
        hits = np.zeros(n, HIT)
        uv = np.zeros((n, 2), np.float32)
        second = np.zeros(n, np.float32)
        cnt = np.zeros((n, 4), np.uint32) if counts else None
        lib().orc_trace_closest(self._h, r.ctypes.data, n, hits.ctypes.data, uv.ctypes.data, second.ctypes.data,
                                None if cnt is None else cnt.ctypes.data, int(cull), threads)
        out = {"hits": hits, "uv": uv, "second_t": second}
        if counts:
            out["counts"] = cnt
        return out

    def render(self, conf: Conf, camera: Camera, iter_count: int = 0, history: np.ndarray | None = None,
               rows: tuple[int, int] | None = None, threads: int = 8, cull: bool = True) -> dict:
        w, h = conf.render_width, conf.render_height
        hist = np.zeros(w * h, HISTORY) if history is None else np.ascontiguousarray(history.reshape(-1), HISTORY).copy()
        rgba = np.zeros((h, w, 4), np.uint8)
        ph = np.zeros(w * h, np.uint32)
        rc = np.zeros(w * h, np.uint32)
        y0, y1 = rows if rows else (0, h)
        st = lib().orc_render(self._h, C.byref(conf), C.byref(camera), iter_count, hist.ctypes.data, rgba.ctypes.data,
                              ph.ctypes.data, rc.ctypes.data, y0, y1, threads, int(cull))
        if st != 0:
            raise RuntimeError("oracle: scene uses features the oracle does not restate")
        return {"history": hist.reshape(h, w), "rgba": rgba, "path_hash": ph.reshape(h, w),
                "ray_count": rc.reshape(h, w)}

    def close(self) -> None:
        if self._h:
            lib().orc_scene_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
