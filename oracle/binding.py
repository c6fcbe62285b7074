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


# Lobe models of the sampling code (sample.wesl:45-52).
DIFFUSE, SPEC_REFLECT, ROUGH_REFLECT, SPEC_TRANSMIT, ROUGH_TRANSMIT = 1, 2, 3, 4, 5


class OrcMat(C.Structure):
    """A decoded material at one wavelength (cg_oracle.h orc_mat)."""
    _fields_ = [("diffuse", C.c_float), ("alpha", C.c_float), ("ior_re", C.c_float), ("ior_im", C.c_float),
                ("transmission", C.c_float), ("volume", C.c_uint32)]


def mat(diffuse=0.0, alpha=0.0, ior=1.5, ior_im=0.0, transmission=0.0, volume=False) -> OrcMat:
    return OrcMat(diffuse, alpha, ior, ior_im, transmission, int(bool(volume)))


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
            ("orc_render_mode", C.c_int, [P, C.POINTER(Conf), C.POINTER(Camera), U32, P, P, P, P, U32, U32,
                                          C.c_int, C.c_int, C.c_int]),
            ("orc_sample_seed", U32, [U32, U32]),
            ("orc_sample_carry", U32, [U32]),
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
            ("orc_lobe_sample", F, [C.c_int, FP, FP, C.POINTER(OrcMat), C.POINTER(U32), FP]),
            ("orc_lobe_eval", F, [C.c_int, FP, FP, C.POINTER(OrcMat), FP, FP]),
            ("orc_ggx_d", F, [FP, FP, F]),
            ("orc_ggx_g1", F, [F, F]),
            ("orc_ggx_g2", F, [F, F, F]),
            ("orc_sample_micro_normal", None, [FP, FP, F, C.POINTER(U32), FP]),
            ("orc_tangent_space", None, [FP, FP, FP]),
            ("orc_reflect", None, [FP, FP, FP]),
            ("orc_refract", None, [FP, FP, F, FP]),
            ("orc_sample_counts", None, [C.POINTER(OrcMat), C.c_int, C.c_int, C.POINTER(Conf), C.POINTER(U32)]),
            ("orc_balance_denominator", F, [P, FP, FP, FP, C.POINTER(OrcMat), C.POINTER(U32), C.c_int, F, F, FP]),
            ("orc_light_pdf", F, [P, FP, FP]),
            ("orc_emission_sample", F, [P, F, FP, C.POINTER(U32), FP, FP, FP, C.POINTER(U32)]),
            ("orc_spawn", None, [FP, FP, FP, FP]),
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


def lobe_sample(model: int, d, n, m: OrcMat, state: int) -> tuple[np.ndarray, float, int]:
    """(direction, pdf, next rng state) of one draw of a lobe."""
    st = U32(state)
    out = (C.c_float * 3)()
    pdf = lib().orc_lobe_sample(model, fvec(d), fvec(n), C.byref(m), C.byref(st), out)
    return np.array(out[:], np.float32), float(pdf), int(st.value)


def lobe_eval(model: int, d, n, m: OrcMat, nxt) -> tuple[float, float]:
    """(pdf, dist) of a lobe at a given next direction."""
    dist = C.c_float()
    pdf = lib().orc_lobe_eval(model, fvec(d), fvec(n), C.byref(m), fvec(nxt), C.byref(dist))
    return float(pdf), float(dist.value)


def vec_out(fn, *args) -> np.ndarray:
    out = (C.c_float * 3)()
    fn(*args, out)
    return np.array(out[:], np.float32)


def sample_counts(m: OrcMat, can_sample: bool, depth: int, conf: Conf) -> tuple[int, ...]:
    out = (U32 * 6)()
    lib().orc_sample_counts(C.byref(m), int(can_sample), depth, C.byref(conf), out)
    return tuple(int(x) for x in out)


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
               rows: tuple[int, int] | None = None, threads: int = 8, cull: bool = True, rng_mode: int = 0) -> dict:
        """rng_mode 0: the reference's per-pixel stream; 1: one stream per sample (orc_render_mode)."""
        w, h = conf.render_width, conf.render_height
        hist = np.zeros(w * h, HISTORY) if history is None else np.ascontiguousarray(history.reshape(-1), HISTORY).copy()
        rgba = np.zeros((h, w, 4), np.uint8)
        ph = np.zeros(w * h, np.uint32)
        rc = np.zeros(w * h, np.uint32)
        y0, y1 = rows if rows else (0, h)
        st = lib().orc_render_mode(self._h, C.byref(conf), C.byref(camera), iter_count, hist.ctypes.data,
                                   rgba.ctypes.data, ph.ctypes.data, rc.ctypes.data, y0, y1, threads, int(cull),
                                   int(rng_mode))
        if st != 0:
            raise RuntimeError("oracle: scene uses features the oracle does not restate")
        return {"history": hist.reshape(h, w), "rgba": rgba, "path_hash": ph.reshape(h, w),
                "ray_count": rc.reshape(h, w)}

    def light_pdf(self, pos, next_dir) -> float:
        return float(lib().orc_light_pdf(self._h, fvec(pos), fvec(next_dir)))

    def emission_sample(self, lam: float, pos, state: int) -> dict:
        st = U32(state)
        d = (C.c_float * 3)()
        dist, rad, prim = C.c_float(), C.c_float(), U32()
        pdf = lib().orc_emission_sample(self._h, lam, fvec(pos), C.byref(st), d, C.byref(dist), C.byref(rad),
                                        C.byref(prim))
        return {"dir": np.array(d[:], np.float32), "pdf": float(pdf), "dist": float(dist.value),
                "radiance": float(rad.value), "prim": int(prim.value), "state": int(st.value)}

    def balance_denominator(self, d, pos, n, m: OrcMat, counts, model: int, model_pdf: float, light_pdf: float,
                            next_dir) -> float:
        c = (U32 * 6)(*counts)
        return float(lib().orc_balance_denominator(self._h, fvec(d), fvec(pos), fvec(n), C.byref(m), c, model,
                                                   model_pdf, light_pdf, fvec(next_dir)))

    def close(self) -> None:
        if self._h:
            lib().orc_scene_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
