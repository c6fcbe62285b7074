"""Pin the oracle (CPU restatement of the reference path) without the
reference running: independent integer models, FP64 closed forms, special
cases and brute force.  The reference holds no golden vectors, tests or
fixtures for this path, so these checks are what fixes the oracle."""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import integrate, special

from helpers import MISS

# --- RNG: math/rng.wesl:1-21 -------------------------------------------------


def pcg_hash_model(s: int) -> int:
    """Independent integer model of the reference's stepRng (Jarzynski & Olano
    2020 PCG-RXS-M-XS hash), in Python's unbounded integers masked to 32 bits."""
    s = (s * 747796405 + 2891336453) % 2**32
    w = (((s >> ((s >> 28) + 4)) ^ s) * 277803737) % 2**32
    return ((w >> 22) ^ w) % 2**32


def test_rng_known_answers(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(0)
    states = [0, 1, 2, 0xFFFFFFFF, 0x80000000, 747796405] + [int(x) for x in rng.integers(0, 2**32, 2000)]
    for s in states:
        assert L.orc_rng_step(s) == pcg_hash_model(s)
    # a 10-step stream from seed 0 (the pixel at (0, 0) on frame 0)
    s, seq = 0, []
    for _ in range(10):
        s = pcg_hash_model(s)
        seq.append(s)
    t = 0
    for v in seq:
        t = L.orc_rng_step(t)
        assert t == v


def test_rand_float_scaling_and_one(oracle):
    L = oracle.lib()
    # randFloat = f32(state) / 2^32; states >= 2^32 - 128 round to exactly 1.0
    st = C.c_uint32(0)
    for _ in range(1000):
        nxt = pcg_hash_model(st.value)
        f = L.orc_rand_float(C.byref(st))
        assert st.value == nxt
        assert f == np.float32(np.float32(nxt) / np.float32(2**32))
    # find a state whose successor is >= 2^32 - 128
    rng = np.random.default_rng(3)
    for _ in range(2_000_000):
        s = int(rng.integers(0, 2**32))
        if pcg_hash_model(s) >= 2**32 - 128:
            st = C.c_uint32(s)
            assert L.orc_rand_float(C.byref(st)) == 1.0
            break


# --- CIE fit: physics/cie.wesl ----------------------------------------------
LOBES = {
    "x": [(0.362, 442.0, 0.0624, 0.0374), (1.056, 599.8, 0.0264, 0.0323), (-0.065, 501.1, 0.0490, 0.0382)],
    "y": [(0.821, 568.8, 0.0213, 0.0247), (0.286, 530.9, 0.0613, 0.0322)],
    "z": [(1.217, 437.0, 0.0845, 0.0278), (0.681, 459.0, 0.0385, 0.0725)],
}


def cmf64(l, lobes):
    return sum(a * math.exp(-0.5 * ((l - c) * (il if l < c else ir)) ** 2) for a, c, il, ir in lobes)


def cmf_integral64(lo, hi, lobes):
    """Closed form of the split-Gaussian integral with the exact erf."""
    tot = 0.0
    for a, c, il, ir in lobes:
        def anti(l):
            w = 1 / il if l < c else 1 / ir
            if l < c:
                return math.sqrt(math.pi / 2) * (1 / il) * (1 + math.erf((l - c) / (math.sqrt(2) * (1 / il))))
            return math.sqrt(math.pi / 2) * (1 / il) + math.sqrt(math.pi / 2) * w * math.erf((l - c) / (math.sqrt(2) * w))
        tot += a * (anti(hi) - anti(lo))
    return tot


def test_erf_against_exact(oracle):
    L = oracle.lib()
    xs = np.linspace(-4, 4, 4001, dtype=np.float32)
    err = max(abs(L.orc_erf(float(x)) - math.erf(float(x))) for x in xs)
    # Abramowitz-Stegun 7.1.26 bound 1.5e-7 plus binary32 rounding
    assert err < 1.5e-7 + 3e-7
    assert L.orc_erf(0.0) == 0.0


def test_cmf_against_fp64(oracle):
    L = oracle.lib()
    out = (C.c_float * 3)()
    for l in np.linspace(380, 720, 341):
        L.orc_xyz_match(float(l), out)
        for k, name in enumerate("xyz"):
            ref = cmf64(float(l), LOBES[name])
            assert abs(out[k] - ref) <= 2e-6 * max(1.0, abs(ref))


def test_cmf_integral_closed_form_and_quadrature(oracle):
    L = oracle.lib()
    total64 = sum(cmf_integral64(380.0, 720.0, LOBES[n]) for n in "xyz")
    quad = sum(integrate.quad(lambda l, n=n: cmf64(l, LOBES[n]), 380, 720, limit=200)[0] for n in "xyz")
    assert abs(total64 - quad) < 1e-8 * total64
    assert abs(total64 - 320.451701239014) < 1e-9
    got = L.orc_xyz_match_sum_integral(380.0, 720.0)
    # fp32 evaluation with the A&S erf: bounded by 7 lobes x (erf error x width) + rounding
    assert abs(got - total64) / total64 < 2e-5
    for hi in (400.0, 500.0, 550.0, 600.0, 700.0):
        ref = sum(cmf_integral64(380.0, hi, LOBES[n]) for n in "xyz")
        assert abs(L.orc_xyz_match_sum_integral(380.0, hi) - ref) < 2e-5 * total64


def test_wavelength_sampler_inverts_cdf(oracle):
    L = oracle.lib()
    total = sum(cmf_integral64(380.0, 720.0, LOBES[n]) for n in "xyz")
    lam, prob = C.c_float(), C.c_float()
    grid = 340.0 / 2**16
    for u in np.linspace(0.0, 1.0, 513):
        L.orc_sample_wavelength(float(u), C.byref(lam), C.byref(prob))
        assert 380.0 <= lam.value <= 720.0
        cdf = sum(cmf_integral64(380.0, lam.value, LOBES[n]) for n in "xyz") / total
        pdf = sum(cmf64(lam.value, LOBES[n]) for n in "xyz") / total
        # the inverse CDF lands within a grid step (plus fp32 CDF error) of u
        assert abs(cdf - min(u, 0.999999)) <= pdf * grid + 3e-5
        assert abs(prob.value - max(pdf, 1e-6)) <= 2e-5 * max(pdf, 1e-3)


def test_illuminate_normalisation(oracle):
    """E[illuminate(l, 1) / p(l)] over the sampler equals (1, 1, 1) x 3 x int(cmf)/total ..."""
    L = oracle.lib()
    xyz = (C.c_float * 3)()
    lam, prob = C.c_float(), C.c_float()
    acc = np.zeros(3)
    n = 20000
    for i in range(n):
        L.orc_sample_wavelength((i + 0.5) / n, C.byref(lam), C.byref(prob))
        L.orc_illuminate(lam.value, 1.0, xyz)
        acc += np.array(xyz[:]) / prob.value
    acc /= n
    # for unit radiance: E = int(xyz) / (total / 3)
    total = sum(cmf_integral64(380.0, 720.0, LOBES[n]) for n in "xyz")
    expect = np.array([cmf_integral64(380.0, 720.0, LOBES[c]) for c in "xyz"]) / (total / 3)
    assert np.allclose(acc, expect, rtol=2e-3)


# --- Smits basis: shading/smits.wesl ----------------------------------------
WHITE = [1.0, 1.0, 0.9999, 0.9993, 0.9992, 0.9998, 1.0, 1.0, 1.0, 1.0]
BLUE = [1.0, 1.0, 0.8916, 0.3323, 0.0, 0.0, 0.0003, 0.0369, 0.0483, 0.0496]


def test_smits_edges_and_quirks(oracle):
    L = oracle.lib()
    assert L.orc_smits(0, 380.0) == np.float32(WHITE[0])
    assert L.orc_smits(3, 720.0) == np.float32(BLUE[9])
    assert L.orc_smits(3, 800.0) == np.float32(BLUE[9])
    # bin = u32(10 (l - 380) / 340): at l = 414 the bin is 1 and the weight
    # extrapolates below BINS[1] = 417.7778 (the reference's 10-for-9 quirk)
    mix = np.float32((np.float32(414.0) - np.float32(417.7778)) / (np.float32(455.5556) - np.float32(417.7778)))
    expect = (np.float32(1) - mix) * np.float32(BLUE[1]) + mix * np.float32(BLUE[2])
    assert L.orc_smits(3, 414.0) == np.float32(expect)
    # l in [686, 720): bin 9 has no upper neighbour -> 0/0 -> NaN
    assert math.isnan(L.orc_smits(0, 700.0))
    # ... which the lifts' max/clamp turn into 0
    rgb = (C.c_float * 3)(0.2, 0.5, 0.9)
    assert L.orc_rgb_to_reflectance(rgb, 700.0) == 0.0
    assert L.orc_rgb_to_radiance(rgb, 700.0) == 0.0


def test_smits_white_lift(oracle):
    L = oracle.lib()
    one = (C.c_float * 3)(1.0, 1.0, 1.0)
    for l in (400.0, 450.0, 500.0, 550.0, 600.0, 650.0):
        w = L.orc_smits(0, l)
        assert L.orc_rgb_to_radiance(one, l) == w
        assert L.orc_rgb_to_reflectance(one, l) == np.float32(min(1.0, np.float32(0.94) * np.float32(w)))


def test_dispersion_at_d_line(oracle):
    L = oracle.lib()
    ld = math.sqrt(523655.0 / 1.5168)  # where the Abbe term vanishes: 587.56 nm
    assert abs(L.orc_dispersed_ior(1.5, 0.30554, ld) - 1.5) < 2e-6
    assert L.orc_dispersed_ior(1.5, 0.30554, 450.0) > 1.5 > L.orc_dispersed_ior(1.5, 0.30554, 650.0)
    assert L.orc_dispersed_ior(1.5, 0.0, 450.0) == np.float32(1.5)


# --- Fresnel: shading/util.wesl ---------------------------------------------


def test_fresnel_closed_forms(oracle):
    L = oracle.lib()
    d = (C.c_float * 3)(0.0, 0.0, -1.0)
    n = (C.c_float * 3)(0.0, 0.0, 1.0)
    for eta in (1.0 / 1.5, 1.5, 1.33):
        f = L.orc_fresnel(d, n, eta, 0.0)
        assert abs(f - ((eta - 1) / (eta + 1)) ** 2) < 1e-6
    # total internal reflection: inside glass at grazing incidence
    g = (C.c_float * 3)(0.9, 0.0, -math.sqrt(1 - 0.81))
    assert L.orc_fresnel(g, n, 1.5, 0.0) == 1.0
    # complex IOR at normal incidence: ((n-1)^2 + k^2) / ((n+1)^2 + k^2)
    for nn, k in ((0.2, 3.0), (1.5, 2.0)):
        f = L.orc_fresnel(d, n, nn, k)
        assert abs(f - ((nn - 1) ** 2 + k * k) / ((nn + 1) ** 2 + k * k)) < 1e-6


# --- Alias table: src/pathtracing/alias_table.h ------------------------------


def _implied(table, n):
    p = np.zeros(n)
    for i, (a, pr) in enumerate(zip(table["alias"], table["prob"])):
        p[i] += pr
        p[a] += 1.0 - pr
    return p / n


@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (7, 2), (4096, 3)])
def test_alias_table_distribution(cg, oracle, n, seed):
    rng = np.random.default_rng(seed)
    w = rng.uniform(0.0, 1.0, n) * 10 ** rng.uniform(0, 2, n)
    out = np.zeros(n, cg.ALIAS)
    oracle.lib().orc_alias_build(w.ctypes.data, n, out.ctypes.data)
    assert np.allclose(_implied(out, n), w / w.sum(), atol=1e-6)
    # the product's host build is the same table bit for bit
    prod = np.zeros(n, cg.ALIAS)
    from crystalgraphics_amd import native
    native.host().cgh_alias_build(w.ctypes.data, n, prod.ctypes.data)
    assert prod.tobytes() == out.tobytes()


def test_alias_table_special_cases(cg, oracle):
    L = oracle.lib()
    w = np.full(5, 2.0)
    out = np.zeros(5, cg.ALIAS)
    L.orc_alias_build(w.ctypes.data, 5, out.ctypes.data)
    assert np.all(out["prob"] == 1.0) and np.all(out["alias"] == np.arange(5))
    z = np.zeros(3)
    out = np.ones(3, cg.ALIAS)
    L.orc_alias_build(z.ctypes.data, 3, out.ctypes.data)
    assert np.all(out["prob"] == 0) and np.all(out["alias"] == 0)
    # sampleEmissivePrim: bin = u32(u1 N), keep it iff u2 < prob[bin]; u1 = 1
    # gives bin N, clamped like the WebGPU robustness clamp
    t = np.zeros(4, cg.ALIAS)
    t["alias"] = [2, 2, 2, 2]
    t["prob"] = [0.5, 1.0, 1.0, 0.25]
    assert L.orc_alias_pick(t.ctypes.data, 4, 0.1, 0.4) == 0
    assert L.orc_alias_pick(t.ctypes.data, 4, 0.1, 0.6) == 2
    assert L.orc_alias_pick(t.ctypes.data, 4, 0.99, 0.1) == 3
    assert L.orc_alias_pick(t.ctypes.data, 4, 0.99, 0.3) == 2


# --- sRGB: physics/srgb.wesl ------------------------------------------------


def test_srgb_closed_form(oracle):
    L = oracle.lib()
    M = np.array([[2.36461385, -0.89654057, -0.46807328], [-0.51516621, 1.4264081, 0.0887581],
                  [0.0052037, -0.01440816, 1.00920446]])
    rgb = (C.c_float * 3)()
    for xyz in ([0.2, 0.3, 0.1], [0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [0.001, 0.0012, 0.0009], [5.0, 5.0, 5.0]):
        L.orc_srgb((C.c_float * 3)(*xyz), rgb)
        lin = M @ np.array(xyz)
        v = np.maximum(lin, 0)
        enc = np.where(v <= 0.0031308, 12.92 * v, 1.055 * v ** (1 / 2.4) - 0.055)
        assert np.allclose(rgb[:], np.clip(enc, 0, 1), atol=2e-6)


# --- Intersection: structural/bvh.wesl:188-221, math/geometric.wesl:5-26 ----


def _ray(cg, o, d):
    r = np.zeros(1, cg.RAY)
    r["org"][0] = o
    r["dir"][0] = d
    return r


def test_triangle_hand_cases(cg, oracle):
    L = oracle.lib()
    tuv = (C.c_float * 3)()
    v0, v1, v2 = (C.c_float * 3)(0, 0, 0), (C.c_float * 3)(1, 0, 0), (C.c_float * 3)(0, 1, 0)
    cases = [  # origin, direction, expect hit, t, u, v
        ((0.25, 0.25, 1.0), (0, 0, -1), True, 1.0, 0.25, 0.25),
        ((0.25, 0.25, -2.0), (0, 0, 1), True, 2.0, 0.25, 0.25),
        ((0.5, 0.5, 1.0), (0, 0, -1), True, 1.0, 0.5, 0.5),     # u + v = 1 edge
        ((0.0, 0.0, 1.0), (0, 0, -1), True, 1.0, 0.0, 0.0),     # vertex
        ((0.6, 0.6, 1.0), (0, 0, -1), False, 0, 0, 0),          # outside
        ((0.25, 0.25, 1.0), (1, 0, 0), False, 0, 0, 0),         # parallel: |a| < 1e-6
        ((0.25, 0.25, -1.0), (0, 0, -1), False, 0, 0, 0),       # behind the origin
        ((0.25, 0.25, 5e-7), (0, 0, -1), False, 0, 0, 0),       # t <= 1e-6
    ]
    for o, d, hit, t, u, v in cases:
        r = _ray(cg, o, d)
        got = L.orc_intersect_triangle(r.ctypes.data, v0, v1, v2, tuv)
        assert bool(got) == hit, (o, d)
        if hit:
            assert tuv[0] == pytest.approx(t, abs=1e-7) and tuv[1] == pytest.approx(u) and tuv[2] == pytest.approx(v)


def test_aabb_hand_cases(cg, oracle):
    L = oracle.lib()
    lb, ub = (C.c_float * 3)(-1, -1, -1), (C.c_float * 3)(1, 1, 1)
    assert L.orc_intersect_aabb(_ray(cg, (0, 0, -5), (0, 0, 1)).ctypes.data, lb, ub)   # zero x/y components
    assert L.orc_intersect_aabb(_ray(cg, (0, 0, 0), (0.3, 0.4, 0.5)).ctypes.data, lb, ub)  # inside
    assert not L.orc_intersect_aabb(_ray(cg, (0, 0, 5), (0, 0, 1)).ctypes.data, lb, ub)  # behind
    assert not L.orc_intersect_aabb(_ray(cg, (3, 0, -5), (0, 0, 1)).ctypes.data, lb, ub)  # beside
    # no t-max: a box beyond any hit still counts (the reference has no culling)
    assert L.orc_intersect_aabb(_ray(cg, (0, 0, -1e5), (0, 0, 1)).ctypes.data, lb, ub)


def _brute_force(desc, rays):
    """Closest hit over every triangle of every instance (no BVH), float32
    Moller-Trumbore in the reference's operation order, vectorised."""
    V = desc["vertices"]["pos"].astype(np.float32)
    I = desc["indices"].reshape(-1, 3)
    best_t = np.full(len(rays), np.float32(1e6))
    second = np.full(len(rays), np.float32(1e6))
    best_inst = np.full(len(rays), MISS, np.uint32)
    best_tri = np.full(len(rays), MISS, np.uint32)
    f32 = np.float32
    for pi, p in enumerate(desc["prims"]):
        inv = p["inv_world"].reshape(4, 4).astype(np.float32)  # rows = columns of the column-major matrix
        o = rays["org"].astype(np.float32)
        d = rays["dir"].astype(np.float32)
        lo = ((inv[0][:3] * o[:, :1] + inv[1][:3] * o[:, 1:2]) + inv[2][:3] * o[:, 2:3]) + inv[3][:3] * f32(1)
        ld = ((inv[0][:3] * d[:, :1] + inv[1][:3] * d[:, 1:2]) + inv[2][:3] * d[:, 2:3]) + inv[3][:3] * f32(0)
        tris = I[p["index_offset"] // 3:(p["index_offset"] + p["index_count"]) // 3]
        for ti, (a, b, c) in enumerate(tris):
            v0, v1, v2 = V[a], V[b], V[c]
            e1, e2 = v1 - v0, v2 - v0
            h = np.stack([ld[:, 1] * e2[2] - ld[:, 2] * e2[1], ld[:, 2] * e2[0] - ld[:, 0] * e2[2],
                          ld[:, 0] * e2[1] - ld[:, 1] * e2[0]], 1)
            aa = (e1[0] * h[:, 0] + e1[1] * h[:, 1]) + e1[2] * h[:, 2]
            ok = ~((aa > f32(-1e-6)) & (aa < f32(1e-6)))
            with np.errstate(divide="ignore", invalid="ignore"):
                f = f32(1) / aa
                s = lo - v0
                u = f * ((s[:, 0] * h[:, 0] + s[:, 1] * h[:, 1]) + s[:, 2] * h[:, 2])
                ok &= ~((u < 0) | (u > 1))
                q = np.stack([s[:, 1] * e1[2] - s[:, 2] * e1[1], s[:, 2] * e1[0] - s[:, 0] * e1[2],
                              s[:, 0] * e1[1] - s[:, 1] * e1[0]], 1)
                v = f * ((ld[:, 0] * q[:, 0] + ld[:, 1] * q[:, 1]) + ld[:, 2] * q[:, 2])
                ok &= ~((v < 0) | (u + v > 1))
                t = f * ((e2[0] * q[:, 0] + e2[1] * q[:, 1]) + e2[2] * q[:, 2])
            ok &= t > f32(1e-6)
            t = np.where(ok, t, f32(1e6))
            second = np.where(t < best_t, best_t, np.minimum(second, t))
            upd = t < best_t
            best_t = np.where(upd, t, best_t)
            best_inst = np.where(upd, pi, best_inst)
            best_tri = np.where(upd, ti, best_tri)
    return best_t, second, best_inst, best_tri


@pytest.mark.parametrize("name,params,n", [("quad", [], 4000), ("spheres", [3], 3000)])
def test_closest_hit_equals_brute_force(cg, oracle, name, params, n):
    s = cg.HostScene.generate(name, seed=21, params=params)
    desc = s.describe()
    rays = s.rays(n, seed=4)
    t, second, inst, tri = _brute_force(desc, rays)
    for cull in (False, True):
        out = oracle.OracleScene.build(desc).trace_closest(rays, cull=cull)
        h = out["hits"]
        tied = (inst != MISS) & ((second - t) <= 1e-6)
        assert np.array_equal(h["inst"][~tied], inst[~tied])
        assert np.array_equal(h["tri"][~tied], tri[~tied])
        assert np.array_equal(h["t"][~tied], t[~tied])
        assert (inst != MISS).sum() > n // 10


# --- BVH build: src/pathtracing/blas.h, tlas.h --------------------------------


def test_blas_invariants(cg, oracle):
    s = cg.HostScene.generate("spheres", seed=8, params=[5])
    desc = s.describe()
    a = oracle.OracleScene.build(desc).arrays()
    nodes, idx = a["blas_nodes"], a["indices"]
    V = desc["vertices"]["pos"]
    seen = np.zeros(len(idx), int)
    for p, inst in zip(desc["prims"], a["instances"]):
        stack = [inst["blas_root_idx"]]
        ntri = p["index_count"] // 3
        while stack:
            n = nodes[stack.pop()]
            if n["triangle_count"] == 0:
                for c in (n["child"], n["child"] + 1):
                    ch = nodes[c]
                    assert np.all(ch["lb"] >= n["lb"] - 1e-6) and np.all(ch["ub"] <= n["ub"] + 1e-6)
                    stack.append(c)
            else:
                assert 1 <= n["triangle_count"] <= 2
                for k in range(n["triangle_count"]):
                    slot = n["child"] + k
                    seen[slot] += 1
                    tv = V[idx[slot]["i"]]
                    assert np.all(tv.min(0) - np.float32(0.001) >= n["lb"] - 1e-6)
                    assert np.all(tv.max(0) + np.float32(0.001) <= n["ub"] + 1e-6)
        del ntri
    assert np.all(seen == 1)


def interior(n):
    return 0 if n <= 2 else 1 + interior(n // 2) + interior(n - n // 2)


def test_blas_node_count_recurrence(cg, oracle):
    for name, params in (("spheres", [4]), ("c0", [])):
        s = cg.HostScene.generate(name, seed=2, params=params)
        desc = s.describe()
        a = oracle.OracleScene.build(desc).arrays()
        expect = sum(1 + 2 * interior(p["index_count"] // 3) for p in desc["prims"])
        assert len(a["blas_nodes"]) == expect
    # the 10M-triangle microbenchmark BLAS: 11,611,391 nodes
    import sys
    sys.setrecursionlimit(10000)
    memo = {}

    def I(n):
        if n <= 2:
            return 0
        if n not in memo:
            memo[n] = 1 + I(n // 2) + I(n - n // 2)
        return memo[n]
    assert 1 + 2 * I(10_000_000) == 11_611_391


# --- Whole-path anchors -------------------------------------------------------


def _empty_scene(cg):
    # one degenerate-free triangle far behind the camera keeps the scene valid
    from crystalgraphics_amd import native
    v = np.zeros(3, cg.VERTEX)
    v["pos"] = [[0, 0, -100], [1, 0, -100], [0, 1, -100]]
    v["nrm"] = [[0, 0, 1]] * 3
    m = np.zeros(1, cg.MATERIAL)
    m["base_color_factor"] = [1, 1, 1, 1]
    m["metallic_factor"] = 0
    m["roughness_factor"] = 1
    m["ior"] = 1.5
    m["tex"]["index"] = -1
    p = np.zeros(1, cg.PRIMITIVE)
    p["vertex_count"] = 3
    p["index_count"] = 3
    p["world"] = np.eye(4, dtype=np.float32).reshape(-1)
    p["inv_world"] = np.eye(4, dtype=np.float32).reshape(-1)
    return cg.HostScene.from_arrays(v, np.arange(3, dtype=np.uint32), m, p)


def test_sky_only_pixels_closed_form(cg, oracle):
    """A pixel whose camera ray escapes sees the sky: per sample,
    XYZ = xyz(l) L_env(l) / (total / 3) / p(l) with p(l) = sum(xyz(l)) / total,
    i.e. 3 xyz(l) L(l) / sum(xyz(l)), computable in FP64 from the same l."""
    s = _empty_scene(cg)
    desc = s.describe()
    orc = oracle.OracleScene.build(desc)
    conf = cg.Conf.make(8, 8, 1, 4)
    cam = cg.Camera.make((0, 0, 0), (1, 0, 0.2), (0.5, 0.5))
    out = orc.render(conf, cam)
    L = oracle.lib()
    for py in range(8):
        for px in range(8):
            seed = px + py * 8
            st = C.c_uint32(pcg_hash_model(seed))  # seedRng steps once
            u = L.orc_rand_float(C.byref(st))
            lam, prob = C.c_float(), C.c_float()
            L.orc_sample_wavelength(u, C.byref(lam), C.byref(prob))
            ray = np.zeros(1, cg.RAY)
            L.orc_camera_ray(C.byref(cam), 8, 8, px, py, ray.ctypes.data)
            d = ray["dir"][0].astype(np.float64)
            d /= np.linalg.norm(d)
            uu = math.atan2(d[1], d[0]) / (2 * math.pi) + 0.5
            vv = math.acos(max(-1, min(1, d[2]))) / math.pi
            y = min(255, int(min(max(1 - vv, 0), 0.999999) * 256))
            t = y / 255.0
            rgb = np.array([0.7, 0.8, 1.0]) * (1 - t) + np.array([0.05, 0.05, 0.06]) * t
            Lr = L.orc_rgb_to_radiance((C.c_float * 3)(*rgb), lam.value)
            xyz = np.array([cmf64(lam.value, LOBES[c]) for c in "xyz"])
            expect = 3 * xyz * Lr / xyz.sum()
            got = out["history"]["xyz"][py, px]
            assert np.allclose(got, expect, rtol=5e-5, atol=1e-7), (px, py)
            del uu


def test_depth_one_is_black_on_hits(cg, oracle):
    s = cg.HostScene.generate("quad", seed=1)
    desc = s.describe()
    orc = oracle.OracleScene.build(desc)
    conf = cg.Conf.make(16, 16, 2, 1)
    out = orc.render(conf, desc["camera"])
    rays = np.zeros(1, cg.RAY)
    hit_any = False
    for py in range(16):
        for px in range(16):
            oracle.lib().orc_camera_ray(C.byref(desc["camera"]), 16, 16, px, py, rays.ctypes.data)
            h = orc.trace_closest(rays)["hits"][0]
            if h["inst"] != MISS:
                hit_any = True
                assert np.all(out["history"]["xyz"][py, px] == 0)
            else:
                assert np.any(out["history"]["xyz"][py, px] != 0)
    assert hit_any


def test_declared_sincos_against_fp64(oracle):
    """The sampling angles' sin/cos (Cody-Waite + Taylor, declared for this
    path) stay within 2 ulp of the exact values over [0, 2 pi]."""
    L = oracle.lib()
    s, c = C.c_float(), C.c_float()
    xs = np.concatenate([np.float32(2 * np.pi) * np.linspace(0, 1, 20001, dtype=np.float32),
                         np.float32([0.0, np.pi / 4, np.pi / 2, np.pi, 3 * np.pi / 2, 2 * np.pi])])
    worst = 0.0
    for x in xs.astype(np.float32):
        L.orc_sincos(float(x), C.byref(s), C.byref(c))
        for got, ref in ((s.value, math.sin(float(x))), (c.value, math.cos(float(x)))):
            ulp = np.spacing(np.float32(max(abs(ref), 2.0 ** -24)))
            worst = max(worst, abs(got - ref) / ulp)
    assert worst <= 2.0


# --- Sample-stream mode (declared for this build) ---------------------------


def sample_seed_model(base: int, i: int) -> int:
    t = pcg_hash_model(base)
    t ^= ((i + 1) * 0x9E3779B9) % 2**32
    return pcg_hash_model(t)


def test_sample_stream_seeds(oracle):
    """Sample i's stream starts at step(step(b) ^ (i + 1) * 0x9E3779B9) and the
    carried seed is step(step(b)), against the independent integer model; the
    seeds of a pixel's samples and its carried seed are pairwise distinct."""
    L = oracle.lib()
    rng = np.random.default_rng(23)
    for base in [0, 1, 2**32 - 1] + [int(x) for x in rng.integers(0, 2**32, 200)]:
        seeds = [L.orc_sample_seed(base, i) for i in range(64)]
        assert seeds == [sample_seed_model(base, i) for i in range(64)]
        carry = L.orc_sample_carry(base)
        assert carry == pcg_hash_model(pcg_hash_model(base))
        assert len(set(seeds + [carry])) == 65


def test_sample_stream_sky_pixels_closed_form(cg, oracle):
    """Sky-only pixels in the sample-stream mode: sample i's wavelength comes
    from the first uniform of its own stream, and the pixel is the in-order
    sum of 3 xyz(l_i) L(l_i) / sum(xyz(l_i)) / spp; the carried seed is
    step(step(b))."""
    s = _empty_scene(cg)
    orc = oracle.OracleScene.build(s.describe())
    spp = 5
    conf = cg.Conf.make(6, 4, spp, 4)
    cam = cg.Camera.make((0, 0, 0), (1, 0, 0.2), (0.5, 0.5))
    out = orc.render(conf, cam, rng_mode=1)
    L = oracle.lib()
    for py in range(4):
        for px in range(6):
            base = px + py * 6
            ray = np.zeros(1, cg.RAY)
            L.orc_camera_ray(C.byref(cam), 6, 4, px, py, ray.ctypes.data)
            d = ray["dir"][0].astype(np.float64)
            d /= np.linalg.norm(d)
            vv = math.acos(max(-1, min(1, d[2]))) / math.pi
            y = min(255, int(min(max(1 - vv, 0), 0.999999) * 256))
            t = y / 255.0
            rgb = np.array([0.7, 0.8, 1.0]) * (1 - t) + np.array([0.05, 0.05, 0.06]) * t
            acc = np.zeros(3, np.float32)
            for i in range(spp):
                st = C.c_uint32(sample_seed_model(base, i))
                u = L.orc_rand_float(C.byref(st))
                lam, prob = C.c_float(), C.c_float()
                L.orc_sample_wavelength(u, C.byref(lam), C.byref(prob))
                Lr = L.orc_rgb_to_radiance((C.c_float * 3)(*rgb), lam.value)
                xyz = np.array([cmf64(lam.value, LOBES[c]) for c in "xyz"])
                acc += np.float32(3 * xyz * Lr / xyz.sum() / spp)
            got = out["history"]["xyz"][py, px]
            assert np.allclose(got, acc, rtol=5e-5, atol=1e-7), (px, py)
            assert out["history"]["rand_seed"][py, px] == pcg_hash_model(pcg_hash_model(base))
