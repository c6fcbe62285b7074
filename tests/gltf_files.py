"""Writers of small glTF 2.0 test files (.gltf with an external buffer and
images, .gltf with data URIs only, .glb), of PNG images with every row
filter and of Radiance RGBE (.hdr) images with flat or run-length rows, plus the scene each file must load as, restated independently of the
library's loader: vertex and index arrays rebased per primitive, the default
material at index 0 and glTF material i at i + 1, primitive objects in
depth-first node order with world = Rx(+90 deg) * parent * local (local =
matrix or T * R * S), texture slots with scale copied into strength.

The reference holds no glTF assets, so these files are the loader's fixtures;
they are generated on the fly from the fixed data below."""
from __future__ import annotations

import base64
import json
import math
import struct
import zlib
from pathlib import Path

import numpy as np

REPEAT, CLAMP, MIRROR = 10497, 33071, 33648


# --- PNG ----------------------------------------------------------------------------------------------------------
def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if pa <= pb and pa <= pc else (b if pb <= pc else c)


def _filter_rows(raw: np.ndarray, bpp: int) -> bytes:
    """Filter each row with type (row index mod 5), so all five filters occur."""
    h, row = raw.shape
    out = bytearray()
    prev = np.zeros(row, np.int64)
    for y in range(h):
        cur = raw[y].astype(np.int64)
        ft = y % 5
        left = np.concatenate([np.zeros(bpp, np.int64), cur[:-bpp]]) if row > bpp else np.zeros(row, np.int64)
        upleft = np.concatenate([np.zeros(bpp, np.int64), prev[:-bpp]]) if row > bpp else np.zeros(row, np.int64)
        if ft == 0:
            f = cur
        elif ft == 1:
            f = cur - left
        elif ft == 2:
            f = cur - prev
        elif ft == 3:
            f = cur - (left + prev) // 2
        else:
            f = cur - np.array([_paeth(int(a), int(b), int(c)) for a, b, c in zip(left, prev, upleft)], np.int64)
        out.append(ft)
        out += (f % 256).astype(np.uint8).tobytes()
        prev = cur
    return bytes(out)


def _chunk(tag: bytes, data: bytes) -> bytes:
    return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)


def png(pixels: np.ndarray, color_type: int, depth: int = 8, palette: np.ndarray | None = None,
        trns: bytes | None = None) -> bytes:
    """pixels: (h, w, channels) integer samples of the given bit depth (palette indices for type 3)."""
    h, w = pixels.shape[:2]
    ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[color_type]
    px = pixels.reshape(h, w * ch).astype(np.int64)
    if depth == 16:
        raw = np.zeros((h, w * ch * 2), np.int64)
        raw[:, 0::2] = px >> 8
        raw[:, 1::2] = px & 255
    elif depth == 8:
        raw = px
    else:  # pack sub-byte samples, most significant first
        per = 8 // depth
        row = (w * ch * depth + 7) // 8
        raw = np.zeros((h, row), np.int64)
        for x in range(w * ch):
            raw[:, x // per] |= px[:, x] << (8 - depth * (x % per + 1))
    bpp = max(1, ch * depth // 8)
    ihdr = struct.pack(">IIBBBBB", w, h, depth, color_type, 0, 0, 0)
    out = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", ihdr)
    if palette is not None:
        out += _chunk(b"PLTE", palette.astype(np.uint8).tobytes())
    if trns is not None:
        out += _chunk(b"tRNS", trns)
    data = zlib.compress(_filter_rows(raw, bpp))
    out += _chunk(b"IDAT", data[: len(data) // 2]) + _chunk(b"IDAT", data[len(data) // 2:])  # split IDAT
    return out + _chunk(b"IEND", b"")


# --- Radiance RGBE (.hdr) -------------------------------------------------------------------------------------------
def _rle_channel(vals: np.ndarray) -> bytes:
    """New-style RLE of one component of a row: runs of >= 3 equal bytes as (128 + n, v), the rest as literals."""
    out, i, n = bytearray(), 0, len(vals)
    while i < n:
        j = i
        while j < n and vals[j] == vals[i] and j - i < 127:
            j += 1
        if j - i >= 3:
            out += bytes([128 + (j - i), int(vals[i])])
            i = j
            continue
        k = i
        while k < n and k - i < 128 and not (k + 2 < n and vals[k] == vals[k + 1] == vals[k + 2]):
            k += 1
        k = max(k, i + 1)
        out += bytes([k - i]) + bytes(int(v) for v in vals[i:k])
        i = k
    return bytes(out)


def hdr_file(rgbe: np.ndarray, rle: bool, magic: bytes = b"#?RADIANCE", extra_header: bytes = b"",
             res_line: bytes | None = None) -> bytes:
    h, w = rgbe.shape[:2]
    head = magic + b"\n# written by the tests\n" + extra_header + b"FORMAT=32-bit_rle_rgbe\nEXPOSURE=1.0\n\n"
    head += res_line if res_line is not None else f"-Y {h} +X {w}\n".encode()
    if not rle:
        return head + rgbe.astype(np.uint8).tobytes()
    body = bytearray()
    for y in range(h):
        body += bytes([2, 2, w >> 8, w & 255])
        for k in range(4):
            body += _rle_channel(rgbe[y, :, k])
    return head + bytes(body)


def images():
    """Four images (encoded PNG, expected RGBA8) covering RGBA8, RGB8, 4-bit palette with alpha, 16-bit gray."""
    rng = np.random.default_rng(7)
    out = []
    a = rng.integers(0, 256, (5, 6, 4))
    out.append((png(a, 6), a.astype(np.uint8)))
    b = rng.integers(0, 256, (4, 3, 3))
    out.append((png(b, 2), np.concatenate([b, np.full((4, 3, 1), 255)], axis=2).astype(np.uint8)))
    pal = rng.integers(0, 256, (16, 3))
    alpha = bytes(rng.integers(0, 256, 10).astype(np.uint8))
    idx = rng.integers(0, 16, (3, 7, 1))
    exp = np.zeros((3, 7, 4), np.uint8)
    exp[..., :3] = pal[idx[..., 0]]
    exp[..., 3] = np.where(idx[..., 0] < 10, np.frombuffer(alpha, np.uint8)[np.minimum(idx[..., 0], 9)], 255)
    out.append((png(idx, 3, 4, pal, alpha), exp))
    g = rng.integers(0, 65536, (2, 9, 1))
    hi = (g >> 8).astype(np.uint8)
    out.append((png(g, 0, 16), np.concatenate([hi, hi, hi, np.full((2, 9, 1), 255, np.uint8)], axis=2)))
    return out


# --- scene --------------------------------------------------------------------------------------------------------
def _quat(axis, ang):
    axis = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    s = math.sin(ang / 2)
    return [float(axis[0] * s), float(axis[1] * s), float(axis[2] * s), math.cos(ang / 2)]


NODES = [
    {"name": "a", "mesh": 0, "children": [1], "translation": [0.5, 0.25, -1.0],
     "rotation": _quat([0.3, 1.0, 0.2], 0.7), "scale": [1.5, 0.75, 1.25]},
    {"name": "b", "mesh": 1,
     "matrix": [0.8, 0.0, -0.6, 0.0, 0.0, 1.0, 0.0, 0.0, 0.6, 0.0, 0.8, 0.0, -0.2, 0.3, 0.4, 1.0]},
    {"name": "c", "children": [3], "translation": [-1.0, 0.0, 0.0]},
    {"name": "d", "mesh": 0, "scale": [0.5, 0.5, 0.5]},
]
SCENE_NODES = [0, 2]


def _local(n):
    if "matrix" in n:
        return np.array(n["matrix"], np.float64).reshape(4, 4).T
    t = n.get("translation", [0, 0, 0])
    x, y, z, w = n.get("rotation", [0, 0, 0, 1])
    s = n.get("scale", [1, 1, 1])
    r = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    m = np.eye(4)
    m[:3, :3] = r @ np.diag(s)
    m[:3, 3] = t
    return m


def _meshes(rng):
    """Two meshes: mesh 0 with two primitives (material 0, and none), mesh 1 with one (material 1)."""
    def grid(n):
        p = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
        nr = rng.normal(size=(n, 3)).astype(np.float32)
        uv0 = rng.uniform(-0.5, 1.5, (n, 2)).astype(np.float32)
        uv1 = rng.integers(0, 256, (n, 2)).astype(np.uint8)  # normalised unsigned bytes
        return p, nr, uv0, uv1
    p0 = grid(6)
    i0 = np.array([0, 1, 2, 2, 3, 0, 3, 4, 5], np.uint8)
    p1 = grid(4)
    i1 = np.array([0, 2, 1, 1, 2, 3], np.uint16)
    p2 = grid(3)
    i2 = np.array([0, 1, 2], np.uint32)
    return [[(p0, i0, 0), (p1, i1, None)], [(p2, i2, 1)]]


MATERIALS = [
    {"name": "textured", "pbrMetallicRoughness": {"baseColorFactor": [0.9, 0.8, 0.7, 1.0], "metallicFactor": 0.0,
                                                  "roughnessFactor": 0.6,
                                                  "baseColorTexture": {"index": 0, "texCoord": 1},
                                                  "metallicRoughnessTexture": {"index": 2}},
     "normalTexture": {"index": 3, "scale": 0.7}, "occlusionTexture": {"index": 3, "strength": 0.5},
     "emissiveTexture": {"index": 1}, "emissiveFactor": [0.2, 0.1, 0.05], "doubleSided": True,
     "alphaMode": "MASK", "alphaCutoff": 0.25},
    {"name": "glass", "pbrMetallicRoughness": {"metallicFactor": 0.0, "roughnessFactor": 0.0},
     "emissiveFactor": [1.0, 0.9, 0.8],
     "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 6.0},
                    "KHR_materials_transmission": {"transmissionFactor": 1.0, "transmissionTexture": {"index": 2}},
                    "KHR_materials_ior": {"ior": 1.45},
                    "KHR_materials_dispersion": {"dispersion": 0.3},
                    "KHR_materials_volume": {"thicknessFactor": 0.5, "attenuationColor": [0.9, 0.95, 1.0]},
                    "KHR_materials_specular": {"specularFactor": 0.5},
                    "KHR_materials_unlit": {}}},
]
SAMPLERS = [{"wrapS": CLAMP, "wrapT": MIRROR}, {}]
TEXTURES = [{"source": 0, "sampler": 0}, {"source": 1, "sampler": 1}, {"source": 2}, {"source": 3, "sampler": 0}]


def write(directory: Path, form: str, extra: dict | None = None) -> Path:
    """form: "gltf" (external .bin and .png files), "datauri" (everything inline), "glb"."""
    rng = np.random.default_rng(3)
    meshes = _meshes(rng)
    blob = bytearray()
    views, accessors = [], []

    def add(arr: np.ndarray, typ: str, comp: int, normalized=False) -> int:
        while len(blob) % 4:
            blob.append(0)
        views.append({"buffer": 0, "byteOffset": len(blob), "byteLength": arr.nbytes})
        blob.extend(arr.tobytes())
        acc = {"bufferView": len(views) - 1, "componentType": comp, "count": int(arr.shape[0]), "type": typ}
        if normalized:
            acc["normalized"] = True
        accessors.append(acc)
        return len(accessors) - 1

    gl_meshes = []
    for prims in meshes:
        gp = []
        for (p, nr, uv0, uv1), idx, mat in prims:
            attrs = {"POSITION": add(p, "VEC3", 5126), "NORMAL": add(nr, "VEC3", 5126),
                     "TEXCOORD_0": add(uv0, "VEC2", 5126), "TEXCOORD_1": add(uv1, "VEC2", 5121, True)}
            comp = {np.uint8: 5121, np.uint16: 5123, np.uint32: 5125}[idx.dtype.type]
            prim = {"attributes": attrs, "indices": add(idx, "SCALAR", comp)}
            if mat is not None:
                prim["material"] = mat
            gp.append(prim)
        gl_meshes.append({"primitives": gp})
    imgs = images()
    gl_images = []
    for k, (data, _) in enumerate(imgs):
        if form == "datauri":
            gl_images.append({"uri": "data:image/png;base64," + base64.b64encode(data).decode()})
        elif form == "glb" or k == 1:  # one image in the buffer in the .gltf form too
            while len(blob) % 4:
                blob.append(0)
            views.append({"buffer": 0, "byteOffset": len(blob), "byteLength": len(data)})
            blob.extend(data)
            gl_images.append({"bufferView": len(views) - 1, "mimeType": "image/png"})
        else:
            name = f"tex {k}.png"  # a space: the URI is percent-encoded
            (directory / name).write_bytes(data)
            gl_images.append({"uri": name.replace(" ", "%20")})
    doc = {"asset": {"version": "2.0"}, "scene": 0, "scenes": [{"nodes": SCENE_NODES}], "nodes": NODES,
           "meshes": gl_meshes, "materials": MATERIALS, "samplers": SAMPLERS, "textures": TEXTURES,
           "images": gl_images, "accessors": accessors, "bufferViews": views,
           "buffers": [{"byteLength": len(blob)}]}
    if extra:
        doc.update(extra)
    if form == "glb":
        js = json.dumps(doc).encode()
        js += b" " * (-len(js) % 4)
        bn = bytes(blob) + b"\0" * (-len(blob) % 4)
        body = struct.pack("<II", len(js), 0x4E4F534A) + js + struct.pack("<II", len(bn), 0x004E4942) + bn
        path = directory / "scene.glb"
        path.write_bytes(struct.pack("<III", 0x46546C67, 2, 12 + len(body)) + body)
        return path
    if form == "datauri":
        doc["buffers"][0]["uri"] = "data:application/octet-stream;base64," + base64.b64encode(bytes(blob)).decode()
    else:
        (directory / "scene data.bin").write_bytes(bytes(blob))
        doc["buffers"][0]["uri"] = "scene%20data.bin"
    path = directory / f"scene_{form}.gltf"
    path.write_text(json.dumps(doc, indent=1))
    return path


def expected():
    """The scene the files above must load as (vertices, indices, prims, materials' key fields, textures)."""
    rng = np.random.default_rng(3)
    meshes = _meshes(rng)
    verts, indices, mesh_prims = [], [], []
    for prims in meshes:
        mp = []
        for (p, nr, uv0, uv1), idx, mat in prims:
            vo = sum(len(v) for v in verts)
            v = np.zeros((len(p), 10), np.float32)
            v[:, 0:3], v[:, 3:6], v[:, 6:8] = p, nr, uv0
            v[:, 8:10] = uv1.astype(np.float32) / np.float32(255.0)
            verts.append(v)
            mp.append((vo, len(p), sum(len(i) for i in indices), len(idx), 0 if mat is None else mat + 1))
            indices.append(idx.astype(np.uint32) + vo)
        mesh_prims.append(mp)
    rx = np.eye(4)
    rx[1:3, 1:3] = [[0.0, -1.0], [1.0, 0.0]]
    prims = []

    def walk(i, parent):
        world = parent @ _local(NODES[i])
        if "mesh" in NODES[i]:
            for p in mesh_prims[NODES[i]["mesh"]]:
                prims.append((p, world))
        for c in NODES[i].get("children", []):
            walk(c, world)
    for n in SCENE_NODES:
        walk(n, rx)
    return {"vertices": np.concatenate(verts), "indices": np.concatenate(indices), "prims": prims,
            "textures": [img for _, img in images()]}
