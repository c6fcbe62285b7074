"""ctypes bindings of the two native libraries and numpy views of the records
of include/cg_scene.h.  Loading fails loudly if a library is missing: there is
no Python fallback for any compute path."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_DIR = Path(__file__).resolve().parent / "lib"

# --- record dtypes (byte-for-byte include/cg_scene.h) -----------------------
VERTEX = np.dtype([("pos", "<f4", 3), ("_p0", "<f4"), ("nrm", "<f4", 3), ("_p1", "<f4"),
                   ("uv0", "<f4", 2), ("uv1", "<f4", 2)])
TEX_INFO = np.dtype([("index", "<i4"), ("tex_coord", "<u4"), ("scale", "<f4"), ("strength", "<f4")])
MATERIAL = np.dtype([
    ("base_color_factor", "<f4", 4), ("emissive_factor", "<f4", 3), ("emissive_strength", "<f4"),
    ("metallic_factor", "<f4"), ("roughness_factor", "<f4"), ("alpha_cutoff", "<f4"), ("alpha_mode", "<u4"),
    ("ior", "<f4"), ("dispersion", "<f4"), ("transmission_factor", "<f4"), ("thickness_factor", "<f4"),
    ("attenuation_color", "<f4", 3), ("attenuation_distance", "<f4"),
    ("anisotropy_strength", "<f4"), ("anisotropy_rotation", "<f4"), ("clearcoat_factor", "<f4"),
    ("clearcoat_roughness_factor", "<f4"), ("iridescence_factor", "<f4"), ("iridescence_ior", "<f4"),
    ("iridescence_thickness_minimum", "<f4"), ("iridescence_thickness_maximum", "<f4"),
    ("sheen_color_factor", "<f4", 3), ("sheen_roughness_factor", "<f4"),
    ("specular_color_factor", "<f4", 3), ("specular_factor", "<f4"),
    ("flags", "<u4"), ("_padding", "<u4", 3), ("tex", TEX_INFO, 17)])
PRIMITIVE = np.dtype([("vertex_offset", "<u4"), ("vertex_count", "<u4"), ("index_offset", "<u4"),
                      ("index_count", "<u4"), ("material_idx", "<u4"), ("_pad", "<u4", 3),
                      ("world", "<f4", 16), ("inv_world", "<f4", 16)])
TLAS_NODE = np.dtype([("lb", "<f4", 3), ("child", "<u4"), ("ub", "<f4", 3), ("instance_idx", "<u4")])
INSTANCE = np.dtype([("inv_trans", "<f4", 16), ("blas_root_idx", "<u4"), ("material_idx", "<u4"), ("_pad", "<u4", 2)])
BLAS_NODE = np.dtype([("lb", "<f4", 3), ("child", "<u4"), ("ub", "<f4", 3), ("triangle_count", "<u4")])
INDEX = np.dtype([("i", "<u4", 3), ("_pad", "<u4")])
EMISSIVE = np.dtype([("v", "<f4", (3, 4)), ("emission", "<f4", 3), ("area", "<f4"), ("weight", "<f4"),
                     ("_pad", "<f4", 3)])
ALIAS = np.dtype([("alias", "<u4"), ("prob", "<f4")])
RAY = np.dtype([("org", "<f4", 3), ("_p0", "<f4"), ("dir", "<f4", 3), ("_p1", "<f4")])
HIT = np.dtype([("t", "<f4"), ("inst", "<u4"), ("tri", "<u4"), ("_pad", "<u4")])
HISTORY = np.dtype([("xyz", "<f4", 3), ("rand_seed", "<u4")])
assert VERTEX.itemsize == 48 and MATERIAL.itemsize == 432 and PRIMITIVE.itemsize == 160
assert TLAS_NODE.itemsize == 32 and INSTANCE.itemsize == 80 and BLAS_NODE.itemsize == 32
assert INDEX.itemsize == 16 and EMISSIVE.itemsize == 80 and ALIAS.itemsize == 8
assert RAY.itemsize == 32 and HIT.itemsize == 16 and HISTORY.itemsize == 16

MISS_ID = 0xFFFFFFFF
MISS_T = np.float32(1e6)


class Camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("_pad0", C.c_float), ("direction", C.c_float * 3),
                ("_pad1", C.c_float), ("viewport", C.c_float * 2), ("_pad2", C.c_float * 2)]

    @classmethod
    def make(cls, position, direction, viewport=(0.5, 0.5)) -> "Camera":
        c = cls()
        c.position[:] = [float(v) for v in position]
        c.direction[:] = [float(v) for v in direction]
        c.viewport[:] = [float(v) for v in viewport]
        return c


class Conf(C.Structure):
    _fields_ = [("window_width", C.c_uint32), ("window_height", C.c_uint32), ("render_width", C.c_uint32),
                ("render_height", C.c_uint32), ("sample_count", C.c_uint32), ("trace_depth", C.c_uint32),
                ("max_transport_sample_count", C.c_uint32), ("max_diffuse_sample_count", C.c_uint32),
                ("max_rough_sample_count", C.c_uint32), ("max_emission_sample_count", C.c_uint32)]

    @classmethod
    def make(cls, render_width=1280, render_height=720, sample_count=1, trace_depth=2,
             max_transport_sample_count=3, max_diffuse_sample_count=1, max_rough_sample_count=1,
             max_emission_sample_count=1) -> "Conf":
        return cls(1920, 1080, render_width, render_height, sample_count, trace_depth,
                   max_transport_sample_count, max_diffuse_sample_count, max_rough_sample_count,
                   max_emission_sample_count)


class SceneDataC(C.Structure):
    """cg_scene_data"""
    _fields_ = [("tlas_nodes", C.c_void_p), ("tlas_node_count", C.c_uint32),
                ("instances", C.c_void_p), ("instance_count", C.c_uint32),
                ("blas_nodes", C.c_void_p), ("blas_node_count", C.c_uint32),
                ("indices", C.c_void_p), ("index_count", C.c_uint32),
                ("tri_ids", C.c_void_p),
                ("vertices", C.c_void_p), ("vertex_count", C.c_uint32),
                ("materials", C.c_void_p), ("material_count", C.c_uint32),
                ("emissive", C.c_void_p), ("emissive_count", C.c_uint32),
                ("alias", C.c_void_p), ("alias_count", C.c_uint32),
                ("env_rgba32f", C.c_void_p), ("env_width", C.c_uint32), ("env_height", C.c_uint32),
                ("tex_rgba8", C.c_void_p), ("tex_width", C.c_uint32), ("tex_height", C.c_uint32),
                ("tex_layers", C.c_uint32)]


class TextureC(C.Structure):
    """cg_texture"""
    _fields_ = [("rgba8", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32), ("wrap_s", C.c_int32),
                ("wrap_t", C.c_int32)]


WRAP_REPEAT, WRAP_CLAMP_TO_EDGE, WRAP_MIRRORED_REPEAT = 10497, 33071, 33648


def texture_structs(textures) -> tuple[C.Array | None, list]:
    """A cg_texture array over a list of {"rgba8": (h, w, 4) uint8, "wrap_s", "wrap_t"}; returns it and the
    arrays to keep alive."""
    if not textures:
        return None, []
    keep = []
    arr = (TextureC * len(textures))()
    for k, t in enumerate(textures):
        img = np.ascontiguousarray(t["rgba8"], np.uint8)
        assert img.ndim == 3 and img.shape[2] == 4, "texture images are (height, width, 4) uint8"
        keep.append(img)
        arr[k] = TextureC(img.ctypes.data, img.shape[1], img.shape[0], int(t.get("wrap_s", WRAP_REPEAT)),
                          int(t.get("wrap_t", WRAP_REPEAT)))
    return arr, keep


class SceneDescC(C.Structure):
    """cgh_scene_desc"""
    _fields_ = [("vertices", C.c_void_p), ("vertex_count", C.c_uint32), ("indices", C.c_void_p),
                ("index_count", C.c_uint32), ("materials", C.c_void_p), ("material_count", C.c_uint32),
                ("prims", C.c_void_p), ("prim_count", C.c_uint32), ("textures", C.c_void_p),
                ("texture_count", C.c_uint32), ("env", C.c_void_p),
                ("env_w", C.c_uint32), ("env_h", C.c_uint32), ("camera", Camera), ("conf", Conf),
                ("tag", C.c_uint64)]


class BatchConf(C.Structure):
    """cg_batch_conf"""
    _fields_ = [("slot_capacity", C.c_uint32), ("max_trace_depth", C.c_uint32), ("rounds_per_launch", C.c_uint32),
                ("use_graph", C.c_uint32), ("max_jobs", C.c_uint32), ("max_stack_depth", C.c_uint32),
                ("block_size", C.c_uint32), ("short_stack", C.c_uint32)]


class Job(C.Structure):
    """cg_job"""
    _fields_ = [("scene", C.c_void_p), ("camera", Camera), ("conf", Conf), ("iter_count", C.c_uint32),
                ("rng_mode", C.c_uint32), ("history", C.c_void_p), ("rgba8", C.c_void_p), ("path_hash", C.c_void_p),
                ("ray_count", C.c_void_p)]


class ResultInfo(C.Structure):
    """cgh_result_info"""
    _fields_ = [("id", C.c_uint64), ("width", C.c_uint32), ("height", C.c_uint32), ("queued_s", C.c_double),
                ("latency_s", C.c_double)]


class ServeStats(C.Structure):
    """cgh_serve_stats"""
    _fields_ = [("submitted", C.c_uint64), ("admitted", C.c_uint64), ("completed", C.c_uint64), ("rays", C.c_uint64),
                ("gpu_busy_s", C.c_double), ("refused", C.c_uint64)]


class Stats(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("nodes", C.c_uint64),
                ("tris", C.c_uint64), ("insts", C.c_uint64)]


SCENE_DATA_ARRAYS = [  # (pointer field, count field, dtype)
    ("tlas_nodes", "tlas_node_count", TLAS_NODE), ("instances", "instance_count", INSTANCE),
    ("blas_nodes", "blas_node_count", BLAS_NODE), ("indices", "index_count", INDEX),
    ("tri_ids", "index_count", np.dtype("<u4")), ("vertices", "vertex_count", VERTEX),
    ("materials", "material_count", MATERIAL), ("emissive", "emissive_count", EMISSIVE),
    ("alias", "alias_count", ALIAS)]


def view_array(ptr: int | None, count: int, dtype: np.dtype) -> np.ndarray:
    """Copy `count` records at `ptr` into a numpy array."""
    if count == 0 or not ptr:
        return np.zeros(0, dtype)
    buf = (C.c_char * (count * dtype.itemsize)).from_address(ptr)
    return np.frombuffer(bytes(buf), dtype=dtype).copy()


def scene_data_arrays(sd: SceneDataC) -> dict:
    out = {}
    for ptr, cnt, dt in SCENE_DATA_ARRAYS:
        out[ptr] = view_array(getattr(sd, ptr), getattr(sd, cnt), dt)
    out["env"] = view_array(sd.env_rgba32f, sd.env_width * sd.env_height * 4, np.dtype("<f4")).reshape(
        sd.env_height, sd.env_width, 4)
    out["tex"] = view_array(sd.tex_rgba8, sd.tex_layers * sd.tex_height * sd.tex_width * 4, np.dtype("u1")).reshape(
        sd.tex_layers, sd.tex_height, sd.tex_width, 4)
    return out


def make_scene_data(arrays: dict) -> tuple[SceneDataC, list]:
    """A cg_scene_data over numpy arrays; returns it and the arrays to keep alive."""
    keep = []
    sd = SceneDataC()
    for ptr, cnt, dt in SCENE_DATA_ARRAYS:
        a = np.ascontiguousarray(arrays[ptr], dtype=dt)
        keep.append(a)
        setattr(sd, ptr, a.ctypes.data if a.size else None)
        if ptr != "tri_ids":
            setattr(sd, cnt, a.shape[0])
    env = np.ascontiguousarray(arrays["env"], dtype=np.float32)
    keep.append(env)
    sd.env_rgba32f = env.ctypes.data
    sd.env_height, sd.env_width = env.shape[0], env.shape[1]
    tex = arrays.get("tex")
    if tex is not None and tex.size:
        tex = np.ascontiguousarray(tex, dtype=np.uint8)
        keep.append(tex)
        sd.tex_rgba8 = tex.ctypes.data
        sd.tex_layers, sd.tex_height, sd.tex_width = tex.shape[0], tex.shape[1], tex.shape[2]
    return sd, keep


def _load(name: str) -> C.CDLL:
    path = LIB_DIR / name
    if not path.exists():
        raise ImportError(f"{path} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
    return C.CDLL(str(path), mode=C.RTLD_GLOBAL)


_dev = None
_host = None

P = C.c_void_p
U32 = C.c_uint32
I32 = C.c_int


def _sig(lib, name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)


def device() -> C.CDLL:
    """libcg_device.so with the signatures of include/cg_device.h."""
    global _dev
    if _dev is None:
        d = _load("libcg_device.so")
        _sig(d, "cg_env_create", I32, C.POINTER(Conf), I32, C.POINTER(P))
        _sig(d, "cg_env_destroy", I32, P)
        _sig(d, "cg_env_last_error", C.c_char_p, P)
        _sig(d, "cg_last_error", C.c_char_p)
        _sig(d, "cg_env_sync", I32, P)
        _sig(d, "cg_env_stream", P, P)
        _sig(d, "cg_env_set_counting", I32, P, I32)
        _sig(d, "cg_env_set_path_records", I32, P, I32)
        _sig(d, "cg_stats_read", I32, P, C.POINTER(Stats))
        _sig(d, "cg_env_set_accel", I32, P, I32)
        _sig(d, "cg_env_set_rng_mode", I32, P, I32)
        _sig(d, "cg_env_set_schedule", I32, P, I32)
        _sig(d, "cg_stats_reset", I32, P)
        _sig(d, "cg_scene_upload", I32, P, C.POINTER(SceneDataC), C.POINTER(P))
        _sig(d, "cg_scene_refit", I32, P, P, C.POINTER(SceneDataC))
        _sig(d, "cg_scene_release", I32, P, P)
        _sig(d, "cg_scene_footprint", I32, P, C.POINTER(C.c_uint64), C.POINTER(U32))
        _sig(d, "cg_layout_check", I32, C.POINTER(SceneDataC), I32, C.POINTER(C.c_uint64), C.POINTER(U32))
        _sig(d, "cg_view", I32, P, P, C.POINTER(Camera), U32)
        _sig(d, "cg_read_target", I32, P, P, P)
        _sig(d, "cg_read_path_records", I32, P, P, P)
        _sig(d, "cg_write_history", I32, P, P)
        _sig(d, "cg_trace_closest", I32, P, P, P, U32, P, P)
        _sig(d, "cg_trace_closest_noalloc", I32, P, P, P, U32, P, P)
        _sig(d, "cg_trace_occluded", I32, P, P, P, P, U32, P)
        _sig(d, "cg_stream_copy", I32, P, P, P, C.c_size_t)
        _sig(d, "cg_stream_read", I32, P, P, C.c_size_t, P)
        _sig(d, "cg_alias_pick", I32, P, P, P, U32, P)
        _sig(d, "cg_device_alloc", I32, P, C.c_size_t, C.POINTER(P))
        _sig(d, "cg_device_free", I32, P, P)
        _sig(d, "cg_memcpy_h2d", I32, P, P, P, C.c_size_t)
        _sig(d, "cg_memcpy_d2h", I32, P, P, P, C.c_size_t)
        _sig(d, "cg_view_job", I32, P, C.POINTER(Job))
        _sig(d, "cg_view_batch_launch", I32, P, C.POINTER(Job), U32, C.POINTER(P))
        _sig(d, "cg_view_batch_query", I32, P, C.POINTER(I32))
        _sig(d, "cg_view_batch_wait", I32, P)
        _sig(d, "cg_view_batch_release", I32, P)
        _sig(d, "cg_batch_create", I32, P, C.POINTER(BatchConf), C.POINTER(P))
        _sig(d, "cg_batch_destroy", I32, P)
        _sig(d, "cg_batch_footprint", I32, P, C.POINTER(C.c_uint64), C.POINTER(U32))
        _sig(d, "cg_batch_submit", I32, P, C.POINTER(Job), C.POINTER(U32))
        _sig(d, "cg_batch_launch", I32, P, U32)
        _sig(d, "cg_batch_poll", I32, P, P, U32, C.POINTER(U32), C.POINTER(U32))
        _sig(d, "cg_batch_drain", I32, P)
        _sig(d, "cg_batch_stats", I32, P, C.POINTER(Stats), I32)
        _dev = d
    return _dev


def host() -> C.CDLL:
    """libcrystal_graphics.so (crystal::graphics host layer)."""
    global _host
    if _host is None:
        device()
        h = _load("libcrystal_graphics.so")
        _sig(h, "cgh_last_error", C.c_char_p)
        _sig(h, "cgh_scene_generate", I32, C.c_char_p, C.c_uint64, P, I32, C.POINTER(P))
        _sig(h, "cgh_scene_create", I32, P, U32, P, U32, P, U32, P, U32, P, U32, P, U32, U32, C.c_uint64,
             C.POINTER(P))
        _sig(h, "cgh_scene_free", None, P)
        _sig(h, "cgh_submit_jobs", I32, P, U32, P, P, P, U32, P, C.POINTER(C.c_double))
        _sig(h, "cgh_scene_describe", I32, P, C.POINTER(SceneDescC))
        _sig(h, "cgh_rays_generate", I32, P, C.c_uint64, U32, P)
        _sig(h, "cgh_scene_data_build", I32, P, I32, I32, C.POINTER(P))
        _sig(h, "cgh_scene_data_view", I32, P, C.POINTER(SceneDataC))
        _sig(h, "cgh_scene_data_free", None, P)
        _sig(h, "cgh_alias_build", I32, P, U32, P)
        _sig(h, "cgh_env_init", I32, C.POINTER(Conf))
        _sig(h, "cgh_env_term", I32)
        _sig(h, "cgh_env_status", I32)
        _sig(h, "cgh_env_poll", I32)
        _sig(h, "cgh_view", I32, P, C.POINTER(Camera))
        _sig(h, "cgh_view_to_buffer", I32, P, C.POINTER(Camera), P, C.c_size_t)
        _sig(h, "cgh_load_scene", I32, C.c_char_p, C.POINTER(P))
        _sig(h, "cgh_serve_last_error", C.c_char_p)
        _sig(h, "cgh_server_create", I32, C.POINTER(P), U32, U32, C.c_uint64, C.c_double, U32, U32, U32, U32, U32,
             U32, U32, C.POINTER(P))
        _sig(h, "cgh_server_free", None, P)
        _sig(h, "cgh_server_submit", I32, P, U32, C.POINTER(Camera), C.POINTER(Conf), C.POINTER(C.c_uint64))
        _sig(h, "cgh_server_next", I32, P, C.c_double, C.POINTER(ResultInfo), P, C.c_size_t)
        _sig(h, "cgh_server_render_alone", I32, P, U32, C.POINTER(Camera), C.POINTER(Conf), C.POINTER(ResultInfo),
             P, C.c_size_t)
        _sig(h, "cgh_server_stats", I32, P, C.POINTER(ServeStats))
        _host = h
    return _host


class NativeError(RuntimeError):
    """A native call returned a non-zero status; carries the code and message."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code
        self.message = message
