"""Pin the shading half of the oracle (the lobes, their sampling, lobe counts,
balance-heuristic MIS, next-event estimation, the camera ray and the spawn
offset) by what the mathematics fixes: published samplers checked with chi^2
tests, FP64 closed forms of the textbook distributions, normalisation
integrals, changes of variables, hand-computed tables and two whole-path
anchors whose expectation is derived by hand.  The reference holds no tests or
fixtures for these functions (it cannot run here), so these checks are what
fixes them; each names the reference lines whose behaviour it pins."""
import collections
import ctypes as C
import math

import numpy as np
import pytest
from scipy import integrate, stats

from helpers import MISS
from test_oracle import LOBES, cmf64, pcg_hash_model

N_UP = np.array([0.0, 0.0, 1.0])


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def basis(n):
    """An orthonormal frame (t, b, n) built here, independent of the oracle's."""
    n = unit(n)
    a = np.array([1.0, 0.0, 0.0]) if abs(n[0]) < 0.6 else np.array([0.0, 1.0, 0.0])
    t = unit(np.cross(a, n))
    return t, np.cross(n, t), n


def fresnel64(cos_i, eta):
    """Unpolarised dielectric Fresnel reflectance, eta = n_incident / n_transmitted."""
    s2 = eta * eta * (1.0 - cos_i * cos_i)
    if s2 >= 1.0:
        return 1.0
    ct = math.sqrt(1.0 - s2)
    rs = (eta * cos_i - ct) / (eta * cos_i + ct)
    rp = (cos_i - eta * ct) / (cos_i + eta * ct)
    return 0.5 * (rs * rs + rp * rp)


def snell64(d, n, eta):
    """Refraction of unit d through a surface with unit normal n facing d's side
    (d . n < 0), eta = n_incident / n_transmitted; None under total reflection."""
    c = -float(np.dot(d, n))
    k = 1.0 - eta * eta * (1.0 - c * c)
    if k < 0:
        return None
    return eta * d + (eta * c - math.sqrt(k)) * n


def draws(oracle, fn, n, seed):
    """n consecutive draws from one RNG stream (as a pixel consumes it)."""
    st = seed
    out = []
    for _ in range(n):
        v, pdf, st = fn(st)
        out.append((v, pdf))
    return out


# --- Lambert: shading/pdf/diffuse.wesl:7-36, dist/diffuse.wesl:6-15 ------------


@pytest.mark.parametrize("normal,incoming", [
    ((0, 0, 1), (0, 0, -1)),
    ((0.3, -0.5, 0.8), (0.5, 0.2, -0.6)),
    ((0.3, -0.5, 0.8), (-0.3, 0.5, 0.9)),  # arriving from below: the lobe is about -n
])
def test_lambert_sampler_is_cosine_weighted(oracle, normal, incoming):
    """Cosine-weighted directions have cos^2(theta) and phi uniform and
    independent about the oriented normal: chi^2 on an 8 x 8 grid."""
    n = unit(normal)
    d = unit(incoming)
    no = -n if np.dot(d, n) > 0 else n
    t, b, _ = basis(no)
    m = oracle.mat(diffuse=0.7, alpha=1.0, ior=1.5)
    N = 64_000
    st = 0x1234567 + int(100 * normal[0])
    u = np.empty(N)
    phi = np.empty(N)
    for i in range(N):
        v, pdf, st = oracle.lobe_sample(oracle.DIFFUSE, d, n, m, st)
        c = float(np.dot(v, no))
        assert c > 0.0
        assert abs(pdf - c / math.pi) <= 1e-5 * max(c / math.pi, 1e-3)
        u[i] = c * c
        phi[i] = math.atan2(float(np.dot(v, b)), float(np.dot(v, t))) % (2 * math.pi)
    h, _, _ = np.histogram2d(u, phi, bins=[8, 8], range=[[0, 1], [0, 2 * math.pi]])
    chi2 = float(((h - N / 64) ** 2 / (N / 64)).sum())
    assert stats.chi2.sf(chi2, 63) > 1e-4, chi2


def test_lambert_eval_closed_form(oracle):
    """pdf = max(cos, 0) / pi and dist = R max(cos, 0) / pi about the normal
    oriented against the incoming direction."""
    rng = np.random.default_rng(3)
    m = oracle.mat(diffuse=0.37, alpha=1.0, ior=1.5)
    for _ in range(2000):
        n = unit(rng.normal(size=3))
        d = unit(rng.normal(size=3))
        w = unit(rng.normal(size=3))
        no = -n if np.dot(d, n) > 0 else n
        c = max(float(np.dot(w, no)), 0.0)
        pdf, dist = oracle.lobe_eval(oracle.DIFFUSE, d, n, m, w)
        assert abs(pdf - c / math.pi) <= 2e-7
        assert abs(dist - 0.37 * c / math.pi) <= 2e-7


# --- Specular reflection: pdf/specular_reflection.wesl, util.wesl ---------------


def test_mirror_direction_and_fresnel(oracle):
    """The sampled direction is the mirror image d - 2 (d.n) n; its pdf is 1 and
    its dist the textbook Fresnel reflectance with the relative IOR of the side
    the ray arrives from; any other direction has pdf 0."""
    rng = np.random.default_rng(5)
    m = oracle.mat(diffuse=0.0, alpha=0.0, ior=1.5)
    for _ in range(500):
        n = unit(rng.normal(size=3))
        d = unit(rng.normal(size=3))
        if abs(np.dot(d, n)) < 0.02:
            continue
        r = d - 2 * np.dot(d, n) * n
        v, pdf, _ = oracle.lobe_sample(oracle.SPEC_REFLECT, d, n, m, 1)
        assert np.abs(v - r).max() < 2e-6 and pdf == 1.0
        assert abs(np.dot(v, n) + np.dot(d, n)) < 2e-6  # angle of reflection = angle of incidence
        p, f = oracle.lobe_eval(oracle.SPEC_REFLECT, d, n, m, r)
        eta = 1 / 1.5 if np.dot(d, n) < 0 else 1.5
        assert p == 1.0 and abs(f - fresnel64(abs(float(np.dot(d, n))), eta)) < 5e-6
        off = unit(r + 0.01 * unit(np.cross(r, rng.normal(size=3))))  # 0.01 rad away
        assert oracle.lobe_eval(oracle.SPEC_REFLECT, d, n, m, off)[0] == 0.0


def test_fresnel_textbook_angles(oracle):
    """Dielectric Fresnel at every angle against the s/p textbook form, both
    sides of the interface (util.wesl:26-60)."""
    L = oracle.lib()
    n = oracle.fvec([0, 0, 1])
    for deg in np.linspace(0, 89.5, 180):
        th = math.radians(deg)
        d = oracle.fvec([math.sin(th), 0, -math.cos(th)])
        for eta in (1 / 1.5, 1.5, 1 / 1.33, 1.33):
            assert abs(L.orc_fresnel(d, n, eta, 0.0) - fresnel64(math.cos(th), eta)) < 3e-6, (deg, eta)


# --- Specular transmission: pdf/specular_transmission.wesl, material.wesl:80-97 -


def test_refraction_obeys_snell(oracle):
    """Volume transmission refracts by Snell's law with the relative IOR of the
    side the ray arrives from (1/n entering, n leaving); past the critical
    angle there is no transmitted direction (pdf 0); a thin surface passes the
    ray straight through; dist = tr (1 - F)."""
    rng = np.random.default_rng(7)
    nd = 1.5
    vol = oracle.mat(alpha=0.0, ior=nd, transmission=0.8, volume=True)
    thin = oracle.mat(alpha=0.0, ior=nd, transmission=0.8, volume=False)
    tir = 0
    for _ in range(600):
        n = unit(rng.normal(size=3))
        d = unit(rng.normal(size=3))
        cd = float(np.dot(d, n))
        if abs(cd) < 0.02:
            continue
        no = n if cd < 0 else -n
        eta = 1 / nd if cd < 0 else nd
        expect = snell64(d, no, eta)
        v, pdf, _ = oracle.lobe_sample(oracle.SPEC_TRANSMIT, d, n, vol, 1)
        if expect is None:
            tir += 1
            assert pdf == 0.0
            continue
        # float32 refraction loses digits as k = 1 - eta^2 sin^2 -> 0 (grazing exit)
        tol = 2e-6 + 1e-7 / math.sqrt(max(1.0 - eta * eta * (1 - cd * cd), 1e-12))
        assert pdf == 1.0 and np.abs(v - expect).max() < tol
        si = float(np.linalg.norm(np.cross(d, no)))
        st_ = float(np.linalg.norm(np.cross(v.astype(np.float64), no)))
        assert abs(st_ - eta * si) < tol  # n1 sin(i) = n2 sin(t)
        assert np.dot(v, no) < 0  # the far side
        assert abs(np.dot(np.cross(d, n), v)) < 5e-6  # plane of incidence
        p, f = oracle.lobe_eval(oracle.SPEC_TRANSMIT, d, n, vol, v)
        assert p == 1.0 and abs(f - 0.8 * (1 - fresnel64(abs(cd), eta))) < 5e-6
        v2, pdf2, _ = oracle.lobe_sample(oracle.SPEC_TRANSMIT, d, n, thin, 1)
        assert pdf2 == 1.0 and np.array_equal(v2, d.astype(np.float32))
    assert tir > 20


def test_wgsl_reflect_refract_builtins(oracle):
    rng = np.random.default_rng(9)
    L = oracle.lib()
    for _ in range(300):
        n = unit(rng.normal(size=3))
        d = unit(rng.normal(size=3))
        if np.dot(d, n) > 0:
            n = -n
        r = oracle.vec_out(L.orc_reflect, oracle.fvec(d), oracle.fvec(n))
        assert np.abs(r - (d - 2 * np.dot(d, n) * n)).max() < 2e-6
        for eta in (0.6, 1.0, 1.4):
            t = oracle.vec_out(L.orc_refract, oracle.fvec(d), oracle.fvec(n), eta)
            e = snell64(d, n, eta)
            k = 1.0 - eta * eta * (1 - float(np.dot(d, n)) ** 2)
            tol = 2e-6 + 1e-7 / math.sqrt(max(k, 1e-12))
            assert np.abs(t - (np.zeros(3) if e is None else e)).max() < tol


# --- GGX: pdf/rough_reflection.wesl:52-139, dist/rough_reflection.wesl ----------


def _ggx_d(oracle, mu, alpha):
    m = oracle.fvec([math.sqrt(max(0.0, 1 - mu * mu)), 0.0, mu])
    return oracle.lib().orc_ggx_d(m, oracle.fvec([0, 0, 1]), alpha)


MU0 = 2e-3  # below mu = 1e-3 the reference clamps cos^2 at 1e-6 (see below)


@pytest.mark.parametrize("alpha", [0.02, 0.1, 0.3, 0.7, 1.0])
def test_ggx_ndf_is_normalised(oracle, alpha):
    """The projected microfacet area integrates to one:
    int D(m) (n.m) dw_m = 2 pi int_0^1 D(mu) mu dmu = 1.  With s = mu^2 the
    textbook mass below mu0 is alpha s0 / (1 - (1 - alpha) s0), so the integral
    over [mu0, 1] is one minus that."""
    f = lambda mu: 2 * math.pi * _ggx_d(oracle, mu, alpha) * mu
    pts = sorted({max(MU0, 1 - k * alpha) for k in (30, 10, 3, 1, 0.3)})
    val, err = integrate.quad(f, MU0, 1.0, points=pts, limit=400, epsabs=1e-10, epsrel=1e-10)
    s0 = MU0 * MU0
    expect = 1.0 - alpha * s0 / (1 - (1 - alpha) * s0)
    assert abs(val - expect) < 3e-6, (val, expect, err)


def test_ggx_ndf_grazing_clamp(oracle):
    """The reference divides by max(cos^2, 1e-6) and max(denominator, 1e-6)
    (rough_reflection.wesl:120-139): for micro-normals within 1e-3 of grazing
    D rises far above the textbook value (up to alpha / 1e-6 at cos = 0),
    which adds 7e-5 (alpha 0.02) to 3.5e-3 (alpha 1) of mass below mu = 1e-3."""
    for alpha in (0.02, 0.3, 1.0):
        assert _ggx_d(oracle, 0.0, alpha) == pytest.approx(alpha / 1e-6, rel=1e-6)
        mu = 3e-4
        textbook = alpha / (math.pi * (alpha * mu * mu + 1 - mu * mu) ** 2)
        assert _ggx_d(oracle, mu, alpha) > 10 * textbook


def test_ggx_ndf_equals_textbook_at_sqrt_alpha(oracle):
    """The reference's D is the textbook GGX (Walter 2007) with
    alpha_D = sqrt(alpha) (it writes alpha where the textbook has alpha^2)."""
    for alpha in (0.01, 0.09, 0.25, 0.64, 1.0):
        a2 = alpha
        for mu in np.linspace(0.05, 1.0, 40):
            t2 = (1 - mu * mu) / (mu * mu)
            expect = a2 / (math.pi * mu ** 4 * (a2 + t2) ** 2)
            assert abs(_ggx_d(oracle, mu, alpha) - expect) <= 2e-6 * expect, (alpha, mu)


def test_smith_masking_closed_forms(oracle):
    """G1 = 2 / (1 + sqrt(1 + alpha_D^2 tan^2)) (Smith GGX), exactly 1 at normal
    incidence; G2 is the height-correlated form 1 / (1 + Lambda_1 + Lambda_2)."""
    L = oracle.lib()
    for alpha in (0.01, 0.2, 0.5, 1.0):
        assert L.orc_ggx_g1(1.0, alpha) == 1.0
        lam = lambda c: (math.sqrt(1 + alpha * (1 - c * c) / (c * c)) - 1) / 2
        for c in np.linspace(0.02, 1.0, 50):
            g1 = 2 / (1 + math.sqrt(1 + alpha * (1 - c * c) / (c * c)))
            assert abs(L.orc_ggx_g1(float(c), alpha) - g1) < 3e-6
            assert abs(L.orc_ggx_g2(float(c), 1.0, alpha) - g1) < 3e-6
            for c2 in (0.1, 0.5, 0.9):
                g2 = 1 / (1 + lam(c) + lam(c2))
                assert abs(L.orc_ggx_g2(float(c), c2, alpha) - g2) < 3e-6


@pytest.mark.parametrize("alpha", [0.1, 0.5, 1.0])
def test_visible_normals_integrate_to_the_projected_area(oracle, alpha):
    """For the same roughness in D and G1 (Heitz 2014, eq. 15):
    int G1(v) max(0, v.m) D(m) dw_m = v.n for every view v.  The oracle supplies
    D(mu) and G1; the geometry and the quadrature are FP64 here."""
    L = oracle.lib()
    x, w = np.polynomial.legendre.leggauss(400)
    mu = MU0 + (1 - MU0) * 0.5 * (x + 1)  # the textbook mass below MU0 is < 4e-6
    wmu = (1 - MU0) * 0.5 * w
    D = np.array([_ggx_d(oracle, float(m_), alpha) for m_ in mu])
    phi = (np.arange(2048) + 0.5) * (2 * math.pi / 2048)
    for deg in (0, 30, 60, 80):
        th = math.radians(deg)
        v = np.array([math.sin(th), 0.0, math.cos(th)])
        g1 = L.orc_ggx_g1(float(v[2]), alpha)
        s = np.sqrt(1 - mu * mu)[:, None]
        vm = v[0] * s * np.cos(phi)[None, :] + v[2] * mu[:, None]
        inner = np.maximum(vm, 0).sum(axis=1) * (2 * math.pi / 2048)
        val = g1 * float((wmu * D * inner).sum())
        assert abs(val - v[2]) < 2e-3 * max(v[2], 0.2), (deg, val, v[2])


def _vndf64(mx, my, mz, v, a):
    """Heitz 2018's distribution of visible normals for the textbook GGX of
    roughness a, view v (unit, v.z > 0), at micro-normals (mx, my, mz)."""
    t2 = (mx * mx + my * my) / (mz * mz)
    D = 1.0 / (math.pi * a * a * mz ** 4 * (1 + t2 / (a * a)) ** 2)
    tv2 = (v[0] ** 2 + v[1] ** 2) / (v[2] ** 2)
    g1 = 2 / (1 + math.sqrt(1 + a * a * tv2))
    return g1 * np.maximum(0, v[0] * mx + v[1] * my + v[2] * mz) * D / v[2]


@pytest.mark.parametrize("alpha,deg", [(0.15, 0), (0.15, 60), (0.5, 35), (0.5, 75)])
def test_visible_normal_sampler_chi2(oracle, alpha, deg):
    """The micro-normal sampler (rough_reflection.wesl:52-103) is Heitz 2018's
    visible-normal sampler with stretch alpha (the sampler's own roughness,
    where D uses sqrt(alpha)): chi^2 of 20,000 draws against the textbook VNDF,
    binned in w = tan^2 / (alpha^2 + tan^2) and azimuth."""
    L = oracle.lib()
    th = math.radians(deg)
    v = np.array([math.sin(th), 0.0, math.cos(th)])
    d = oracle.fvec(-v)
    nvec = oracle.fvec([0, 0, 1])
    N = 20_000
    st = C.c_uint32(987654 + deg)
    out = (C.c_float * 3)()
    ws = np.empty(N)
    ph = np.empty(N)
    for i in range(N):
        L.orc_sample_micro_normal(d, nvec, alpha, C.byref(st), out)
        mx, my, mz = out[0], out[1], out[2]
        t2 = (mx * mx + my * my) / (mz * mz)
        ws[i] = t2 / (alpha * alpha + t2)
        ph[i] = math.atan2(my, mx) % (2 * math.pi)
    nw, nph = 10, 12
    obs, _, _ = np.histogram2d(ws, ph, bins=[nw, nph], range=[[0, 1], [0, 2 * math.pi]])
    # expected bin masses: Gauss-Legendre in (w, phi) with the Jacobian
    # dw_m = dmu dphi, mu = 1 / sqrt(1 + T), T = alpha^2 w / (1 - w)
    gx, gw = np.polynomial.legendre.leggauss(48)  # bins hold the kink of max(0, v.m)
    exp = np.zeros((nw, nph))
    for i in range(nw):
        # Gauss-Legendre in q = sqrt(1 - w): the density in w has a 1/sqrt(1 - w)
        # singularity at w = 1 (grazing micro-normals), which dw = 2 q dq removes
        w0, w1 = i / nw, (i + 1) / nw
        q0, q1 = math.sqrt(1 - w1), math.sqrt(1 - w0)
        q = q0 + (q1 - q0) * 0.5 * (gx + 1)
        wq = 1 - q * q
        wwq = (q1 - q0) * 0.5 * gw * 2 * q
        T = alpha * alpha * wq / (1 - wq)
        mu = 1 / np.sqrt(1 + T)
        jac = 0.5 * (1 + T) ** -1.5 * alpha * alpha / (1 - wq) ** 2
        for j in range(nph):
            p0, p1 = j * 2 * math.pi / nph, (j + 1) * 2 * math.pi / nph
            pq = p0 + (p1 - p0) * 0.5 * (gx + 1)
            wpq = (p1 - p0) * 0.5 * gw
            s = np.sqrt(1 - mu * mu)[:, None]
            val = _vndf64(s * np.cos(pq)[None, :], s * np.sin(pq)[None, :], mu[:, None], v, alpha)
            exp[i, j] = float((wwq[:, None] * wpq[None, :] * jac[:, None] * val).sum())
    assert abs(exp.sum() - 1) < 5e-4, exp.sum()
    e = exp * N
    keep = e > 5
    chi2 = float(((obs[keep] - e[keep]) ** 2 / e[keep]).sum())
    assert obs[~keep].sum() <= max(30, 3 * e[~keep].sum()), "draws in bins the distribution leaves (nearly) empty"
    assert stats.chi2.sf(chi2, int(keep.sum()) - 1) > 1e-4, chi2


@pytest.mark.parametrize("alpha", [0.05, 0.2, 0.6])
def test_rough_reflection_pdf_mass_at_normal_incidence(oracle, alpha):
    """At normal incidence the reflection pdf over outgoing directions is the
    visible-normal pdf times the half-vector Jacobian 1/(4 v.m); it integrates
    over the upper hemisphere to the D-weighted mass of micro-normals within
    45 degrees of n, i.e. 1 / (1 + alpha_D^2) = 1 / (1 + alpha)."""
    m = oracle.mat(alpha=alpha, ior=1.5)
    d = [0.0, 0.0, -1.0]
    f = lambda mu: 2 * math.pi * oracle.lobe_eval(oracle.ROUGH_REFLECT, d, N_UP,
                                                    m, [math.sqrt(max(0.0, 1 - mu * mu)), 0.0, mu])[0]
    pts = sorted({max(0.0, 1 - k * alpha) for k in (30, 10, 3, 1, 0.3)})
    val, _ = integrate.quad(f, 0.0, 1.0, points=pts, limit=400, epsabs=1e-9, epsrel=1e-9)
    assert abs(val - 1 / (1 + alpha)) < 2e-4 * (1 / (1 + alpha)), (val, 1 / (1 + alpha))


def test_rough_reflection_pdf_mass_oblique(oracle):
    """At oblique incidence the same change of variables: the pdf integrated
    over outgoing directions above the surface equals the visible-normal mass
    (textbook, alpha_D = sqrt(alpha)) of micro-normals whose mirror image of
    the view stays above the surface."""
    alpha, deg = 0.3, 60.0
    th = math.radians(deg)
    v = np.array([math.sin(th), 0.0, math.cos(th)])
    m = oracle.mat(alpha=alpha, ior=1.5)
    gx, gw = np.polynomial.legendre.leggauss(160)
    mu = 0.5 * (gx + 1)
    wmu = 0.5 * gw
    nph = 256
    phi = (np.arange(nph) + 0.5) * 2 * math.pi / nph
    lhs = 0.0
    for i, u in enumerate(mu):
        s = math.sqrt(1 - u * u)
        row = sum(oracle.lobe_eval(oracle.ROUGH_REFLECT, -v, N_UP, m, [s * math.cos(p), s * math.sin(p), u])[0]
                  for p in phi)
        lhs += wmu[i] * row * 2 * math.pi / nph
    aD = math.sqrt(alpha)
    gx2, gw2 = np.polynomial.legendre.leggauss(400)
    mu2 = 0.5 * (gx2 + 1)
    w2 = 0.5 * gw2
    ph2 = (np.arange(1024) + 0.5) * 2 * math.pi / 1024
    s2 = np.sqrt(1 - mu2 * mu2)[:, None]
    mx, my, mz = s2 * np.cos(ph2)[None, :], s2 * np.sin(ph2)[None, :], mu2[:, None] * np.ones_like(ph2)[None, :]
    vm = v[0] * mx + v[2] * mz
    oz = 2 * vm * mz - v[2]  # z of the mirror image of v about m
    dens = _vndf64(mx, my, mz, v, aD) * (oz > 0)
    rhs = float((w2[:, None] * dens).sum() * 2 * math.pi / 1024)
    assert abs(lhs - rhs) < 5e-3 * rhs, (lhs, rhs)


@pytest.mark.parametrize("alpha", [0.05, 0.3])
def test_rough_transmission_pdf_change_of_variables(oracle, alpha):
    """At normal incidence into glass (n = 1.5) the transmission pdf
    (rough_transmission.wesl:129-181) is vnp(m) (i.m) / den^2 with
    den = o.m + (i.m)/n; with the refraction Jacobian dw_m/dw_o = |o.m| / den^2
    its integral over outgoing directions equals int vnp(m) (i.m) / |o(m).m| dw_m,
    where o(m) is the Snell refraction of the view through m: both sides in
    FP64 quadrature, the left from the oracle's pdf, the right from the
    textbook GGX at alpha_D = sqrt(alpha) and Snell's law."""
    nd = 1.5
    m = oracle.mat(alpha=alpha, ior=nd, transmission=1.0)
    d = np.array([0.0, 0.0, -1.0])
    f = lambda mu: 2 * math.pi * oracle.lobe_eval(oracle.ROUGH_TRANSMIT, d, N_UP, m,
                                                    [math.sqrt(max(0.0, 1 - mu * mu)), 0.0, -mu])[0]
    cone = 1 / nd  # refraction from air into n = 1.5 reaches only cos(theta_o) >= 1/n
    pts = sorted({max(cone, 1 - k * alpha) for k in (40, 10, 3, 1, 0.3)})
    lhs, _ = integrate.quad(f, cone, 1.0, points=pts, limit=400, epsabs=1e-10, epsrel=1e-10)
    # Outside that cone the half-vector d/n - o points down and the reference
    # re-orients it (orientedNormal, rough_transmission.wesl:46-56), so the pdf
    # there is positive although no micro-normal refracts into it: that mass
    # is left out of the sampler's distribution.
    beyond, _ = integrate.quad(f, 0.0, cone, limit=400)
    assert beyond > 0.01 * alpha
    st = 55
    for _ in range(3000):
        v, pdf, st = oracle.lobe_sample(oracle.ROUGH_TRANSMIT, d, N_UP, m, st)
        assert pdf == 0.0 or -v[2] >= cone - 1e-6
    a2 = alpha  # alpha_D^2

    def g(mu):
        t2 = (1 - mu * mu) / (mu * mu)
        D = a2 / (math.pi * mu ** 4 * (a2 + t2) ** 2)
        mvec = np.array([math.sqrt(1 - mu * mu), 0.0, mu])
        o = snell64(d, mvec, 1 / nd)
        if o is None or o[2] >= 0:
            return 0.0
        vnp = D * mu  # G1 at normal incidence is 1; (i.m)/(i.n) = mu
        return 2 * math.pi * vnp * mu / abs(float(np.dot(o, mvec)))

    rhs, _ = integrate.quad(g, MU0, 1.0, points=pts, limit=400, epsabs=1e-10, epsrel=1e-10)
    assert abs(lhs - rhs) < 2e-5 * rhs, (lhs, rhs)


def test_rough_lobes_tend_to_the_specular_ones(oracle):
    """As alpha -> 0 rough reflection samples the mirror direction with weight
    dist/pdf -> F (textbook Fresnel), and rough transmission samples the
    Snell direction with weight -> tr (1 - F) at normal incidence."""
    alpha = 1e-4
    refl = oracle.mat(alpha=alpha, ior=1.5)
    for deg in (0.0, 25.0, 50.0, 70.0):
        th = math.radians(deg)
        d = np.array([math.sin(th), 0.0, -math.cos(th)])
        mirror = d - 2 * np.dot(d, N_UP) * N_UP
        st = 4242
        for _ in range(200):
            v, pdf, st = oracle.lobe_sample(oracle.ROUGH_REFLECT, d, N_UP, refl, st)
            assert math.acos(min(1.0, float(np.dot(v, mirror)))) < 5e-3
            p, dist = oracle.lobe_eval(oracle.ROUGH_REFLECT, d, N_UP, refl, v)
            assert abs(dist / p - fresnel64(math.cos(th), 1 / 1.5)) < 2e-3
    trans = oracle.mat(alpha=alpha, ior=1.5, transmission=0.9)
    for deg in (0.0, 30.0, 60.0):
        th = math.radians(deg)
        d = np.array([math.sin(th), 0.0, -math.cos(th)])
        snell = snell64(d, N_UP, 1 / 1.5)
        st = 99
        for _ in range(200):
            v, pdf, st = oracle.lobe_sample(oracle.ROUGH_TRANSMIT, d, N_UP, trans, st)
            assert pdf > 0 and math.acos(min(1.0, float(np.dot(v, snell)))) < 5e-3
            if deg == 0.0:
                p, dist = oracle.lobe_eval(oracle.ROUGH_TRANSMIT, d, N_UP, trans, v)
                assert abs(dist / p - 0.9 * (1 - 0.04)) < 2e-3


# --- Lobe counts: shading/sample.wesl:110-246, lod.wesl:16-18 -------------------

# Decoded materials (diffuse reflectance, alpha, IOR, transmission) and the
# heuristics they give: reflection F0 = ((n-1)^2 + k^2) / ((n+1)^2 + k^2).
MATS = {
    "diffuse": dict(diffuse=0.5, alpha=1.0, ior=1.5),            # D 0.5, rough R 0.04
    "diffuse_ior1": dict(diffuse=0.5, alpha=1.0, ior=1.0),       # D 0.5 only (F0 = 0)
    "mirror_metal": dict(alpha=0.0, ior=0.2, ior_im=3.0),       # specular R 0.923
    "rough_metal": dict(alpha=0.3, ior=0.2, ior_im=3.0),        # rough R 0.923
    "glass": dict(alpha=0.0, ior=1.5, transmission=1.0),        # specular R 0.04, specular T 1
    "rough_glass": dict(alpha=0.2, ior=1.5, transmission=1.0),  # rough R 0.04, rough T 1
    "black": dict(alpha=1.0, ior=1.0),                           # nothing to sample
    "faint_trans": dict(diffuse=0.5, alpha=1.0, ior=1.5, transmission=5e-7),  # T <= 1e-6 ignored
    "mix": dict(diffuse=0.3, alpha=0.0, ior=1.5, transmission=0.5),  # D 0.3, spec R 0.04, spec T 0.5
}

# Transport samples available at lod depth d: u32((D - d) / D * min(T, 4))
# in f32; light samples u32((D - d) / D * min(E, 1)).  Worked by hand:
#   D=8, T=3: d=0 -> 3; d=1 -> 2.625 -> 2; d=2 -> 2.25 -> 2; d=3 -> 1.875 -> 1;
#             d=4 -> 1.5 -> 1; d=5 -> 1.125 -> 1; d=6 -> 0.75 -> 0; d=7 -> 0
#   D=3, T=3: d=1 -> f32(2/3) = 0.666666687 * 3 = 2.00000006 -> f32 2.0 -> 2
#   D=6, T=3: d=4 -> f32(1/3) = 0.333333343 * 3 = 1.00000003 -> f32 1.0 -> 1
#   D=10, T=3: d=7 -> f32(0.3) = 0.300000012 * 3 = 0.90000004 -> 0
# Each lobe with heuristic h > 1e-6 gets min(max(1, u32(h * avail / sum)), cap)
# with cap 1 for every lobe here (DIFFUSE/ROUGH counts are min(conf, 1)); then
# the counts are taken in lobe order (D, specular R, rough R, specular T,
# rough T) until `avail` runs out.  light = 1 iff an emitter exists and the
# light samples available are > 0.
# (material, D, depth, T, DIF, ROUGH, E, can_sample) -> (D, sR, rR, sT, rT, L)
COUNT_TABLE = [
    ("diffuse", 8, 0, 3, 1, 1, 1, True, (1, 0, 1, 0, 0, 1)),
    ("diffuse", 8, 0, 3, 1, 1, 1, False, (1, 0, 1, 0, 0, 0)),
    ("diffuse", 8, 1, 3, 1, 1, 1, True, (1, 0, 1, 0, 0, 0)),   # avail 2, light 0
    ("diffuse", 8, 3, 3, 1, 1, 1, True, (1, 0, 0, 0, 0, 0)),   # avail 1: diffuse first
    ("diffuse", 8, 5, 3, 1, 1, 1, True, (1, 0, 0, 0, 0, 0)),
    ("diffuse", 8, 6, 3, 1, 1, 1, True, (0, 0, 0, 0, 0, 0)),   # avail 0
    ("diffuse", 8, 7, 3, 1, 1, 1, True, (0, 0, 0, 0, 0, 0)),
    ("diffuse", 2, 0, 3, 1, 1, 1, True, (1, 0, 1, 0, 0, 1)),
    ("diffuse", 2, 1, 3, 1, 1, 1, True, (1, 0, 0, 0, 0, 0)),   # 0.5 * 3 = 1.5 -> 1
    ("diffuse", 3, 1, 3, 1, 1, 1, True, (1, 0, 1, 0, 0, 0)),   # 2.00000006 -> 2
    ("diffuse", 6, 4, 3, 1, 1, 1, True, (1, 0, 0, 0, 0, 0)),   # 1.00000003 -> 1
    ("diffuse", 10, 7, 3, 1, 1, 1, True, (0, 0, 0, 0, 0, 0)),  # 0.90000004 -> 0
    ("diffuse", 8, 0, 3, 0, 1, 1, True, (0, 0, 1, 0, 0, 1)),   # DIF 0 caps diffuse at 0
    ("diffuse", 8, 0, 3, 1, 0, 1, True, (1, 0, 0, 0, 0, 1)),   # ROUGH 0 caps rough at 0
    ("diffuse", 8, 0, 3, 1, 1, 0, True, (1, 0, 1, 0, 0, 0)),   # E 0: no light sample
    ("diffuse", 8, 0, 3, 1, 1, 2, True, (1, 0, 1, 0, 0, 1)),   # E capped at 1
    ("diffuse", 8, 0, 9, 1, 1, 1, True, (1, 0, 1, 0, 0, 1)),   # T capped at 4, each lobe at 1
    ("diffuse", 8, 0, 1, 1, 1, 1, True, (1, 0, 0, 0, 0, 1)),
    ("diffuse_ior1", 8, 0, 3, 1, 1, 1, True, (1, 0, 0, 0, 0, 1)),
    ("mirror_metal", 8, 0, 3, 1, 1, 1, True, (0, 1, 0, 0, 0, 1)),
    ("rough_metal", 8, 0, 3, 1, 1, 1, True, (0, 0, 1, 0, 0, 1)),
    ("glass", 8, 0, 3, 1, 1, 1, True, (0, 1, 0, 1, 0, 1)),
    ("glass", 8, 3, 3, 1, 1, 1, True, (0, 1, 0, 0, 0, 0)),     # avail 1: reflection before transmission
    ("rough_glass", 8, 0, 3, 1, 1, 1, True, (0, 0, 1, 0, 1, 1)),
    ("rough_glass", 8, 4, 3, 1, 1, 1, True, (0, 0, 1, 0, 0, 0)),
    ("rough_glass", 8, 0, 3, 1, 0, 1, True, (0, 0, 0, 0, 0, 1)),  # ROUGH 0: both rough lobes capped
    ("black", 8, 0, 3, 1, 1, 1, True, (0, 0, 0, 0, 0, 1)),     # sum 0: no transport, light kept
    ("faint_trans", 8, 0, 3, 1, 1, 1, True, (1, 0, 1, 0, 0, 1)),
    ("mix", 8, 0, 3, 1, 1, 1, True, (1, 1, 0, 1, 0, 1)),
    ("mix", 8, 1, 3, 1, 1, 1, True, (1, 1, 0, 0, 0, 0)),       # avail 2
    ("mix", 8, 3, 3, 1, 1, 1, True, (1, 0, 0, 0, 0, 0)),       # avail 1: diffuse takes it
    ("mix", 8, 0, 2, 1, 1, 1, True, (1, 1, 0, 0, 0, 1)),       # T 2: transmission left out
    ("mix", 8, 0, 2, 0, 1, 1, True, (0, 1, 0, 1, 0, 1)),       # DIF 0 frees a slot for T
]


@pytest.mark.parametrize("row", COUNT_TABLE, ids=[f"{r[0]}-D{r[1]}-d{r[2]}-T{r[3]}-{i}" for i, r in enumerate(COUNT_TABLE)])
def test_lobe_counts_hand_table(cg, oracle, row):
    name, D, depth, T, DIF, ROUGH, E, can, expect = row
    conf = cg.Conf.make(8, 8, 1, D, T, DIF, ROUGH, E)
    assert oracle.sample_counts(oracle.mat(**MATS[name]), can, depth, conf) == expect


# --- Scenes with one emissive triangle ----------------------------------------

LIGHT = np.array([[-0.3, -0.4, 1.0], [0.0, 0.5, 1.0], [0.5, -0.2, 1.0]])


def _light_scene(cg, albedo=0.8, sky=(0.0, 0.0, 0.0), with_light=True, second_light=None):
    """A 4 x 4 m ground quad at z = 0 (diffuse, IOR 1 so no specular lobe) and
    one emissive triangle at z = 1 (black, emissive (1,1,1), strength 1);
    optionally a second emissive triangle; a constant-colour sky."""
    pos = [[-2, -2, 0], [2, -2, 0], [2, 2, 0], [-2, 2, 0]]
    idx = [0, 1, 2, 0, 2, 3]
    lights = ([LIGHT] if with_light else []) + ([second_light] if second_light is not None else [])
    for L in lights:
        idx += [len(pos), len(pos) + 1, len(pos) + 2]
        pos += [list(p) for p in L]
    v = np.zeros(len(pos), cg.VERTEX)
    v["pos"] = pos
    v["nrm"] = [[0, 0, 1]] * 4 + [[0, 0, -1]] * (len(pos) - 4)
    m = np.zeros(2, cg.MATERIAL)
    m["base_color_factor"] = [[albedo, albedo, albedo, 1], [0, 0, 0, 1]]
    m["metallic_factor"] = 0
    m["roughness_factor"] = 1
    m["ior"] = 1.0
    m["tex"]["index"] = -1
    m["emissive_factor"][1] = [1, 1, 1]
    m["emissive_strength"] = [1, 1]
    p = np.zeros(1 + len(lights), cg.PRIMITIVE)
    p["vertex_offset"] = [0] + [4 + 3 * i for i in range(len(lights))]
    p["vertex_count"] = [4] + [3] * len(lights)
    p["index_offset"] = [0] + [6 + 3 * i for i in range(len(lights))]
    p["index_count"] = [6] + [3] * len(lights)
    p["material_idx"] = [0] + [1] * len(lights)
    p["world"] = np.eye(4, dtype=np.float32).reshape(-1)
    p["inv_world"] = np.eye(4, dtype=np.float32).reshape(-1)
    env = np.zeros((1, 1, 4), np.float32)
    env[0, 0, :3] = sky
    env[0, 0, 3] = 1
    return cg.HostScene.from_arrays(v, np.array(idx, np.uint32), m, p, env=env)


def _tri_area(T):
    return 0.5 * np.linalg.norm(np.cross(T[1] - T[0], T[2] - T[0]))


def _ray_tri(o, d, T):
    """FP64 ray-triangle distance (inf if missed)."""
    e1, e2 = T[1] - T[0], T[2] - T[0]
    p = np.cross(d, e2)
    det = np.dot(e1, p)
    if abs(det) < 1e-12:
        return math.inf
    tv = o - T[0]
    u = np.dot(tv, p) / det
    q = np.cross(tv, e1)
    w = np.dot(d, q) / det
    if u < 0 or w < 0 or u + w > 1:
        return math.inf
    t = np.dot(e2, q) / det
    return t if t > 0 else math.inf


@pytest.fixture(scope="module")
def light_scene(cg, oracle):
    s = _light_scene(cg)
    return s, oracle.OracleScene.build(s.describe())


def test_light_pdf_is_area_pdf_over_solid_angle(light_scene, oracle):
    """lightPdf (emission.wesl:120-149) for one emitter of normalised weight 1:
    pdf = (1/A) r^2 / |cos_l| toward points of the triangle, 0 elsewhere."""
    _, orc = light_scene
    A = _tri_area(LIGHT)
    nl = unit(np.cross(LIGHT[1] - LIGHT[0], LIGHT[2] - LIGHT[0]))
    rng = np.random.default_rng(11)
    for _ in range(400):
        x = np.array([rng.uniform(-1.5, 1.5), rng.uniform(-1.5, 1.5), 0.0])
        b = 0.05 + 0.85 * rng.dirichlet([1, 1, 1])  # interior points: no float ties at an edge
        y = b @ LIGHT
        w = unit(y - x)
        r = np.linalg.norm(y - x)
        expect = (1 / A) * r * r / abs(np.dot(nl, w))
        got = orc.light_pdf(x, w)
        assert abs(got - expect) <= 3e-5 * expect
        miss = unit(w + 0.6 * unit(np.cross(w, nl)))
        if _ray_tri(x, miss, LIGHT) == math.inf:
            assert orc.light_pdf(x, miss) == 0.0


def test_light_pdf_normalises_over_directions(light_scene, oracle):
    """int lightPdf dw = 1: directions uniform in a cone around the light from a
    floor point, 100,000 of them, with the cone's solid angle as the weight."""
    _, orc = light_scene
    x = np.array([0.3, -0.6, 0.0])
    axis = unit(LIGHT.mean(axis=0) - x)
    cmax = min(float(np.dot(unit(p - x), axis)) for p in LIGHT) - 0.02
    omega = 2 * math.pi * (1 - cmax)
    t, b, _ = basis(axis)
    rng = np.random.default_rng(13)
    N = 100_000
    c = 1 - rng.random(N) * (1 - cmax)
    phi = rng.random(N) * 2 * math.pi
    s = np.sqrt(1 - c * c)
    dirs = (s * np.cos(phi))[:, None] * t + (s * np.sin(phi))[:, None] * b + c[:, None] * axis
    vals = np.array([orc.light_pdf(x, w) for w in dirs])
    est = omega * vals.mean()
    sem = omega * vals.std() / math.sqrt(N)
    assert abs(est - 1.0) < 5 * sem + 1e-4, (est, sem)


def test_nee_points_uniform_and_pdf_consistent(light_scene, oracle):
    """emissionSample (emission.wesl:46-96): light points are uniform over the
    triangle (chi^2 over its 16 congruent sub-triangles), lie on it, the
    returned distance is |y - x|, and the returned solid-angle pdf is the
    light pdf of the returned direction."""
    _, orc = light_scene
    x = np.array([-0.4, 0.7, 0.0])
    st = 777
    N = 32_000
    keys = collections.Counter()
    e1, e2 = LIGHT[1] - LIGHT[0], LIGHT[2] - LIGHT[0]
    G = np.array([[e1 @ e1, e1 @ e2], [e1 @ e2, e2 @ e2]])
    for i in range(N):
        s = orc.emission_sample(550.0, x, st)
        st = s["state"]
        y = x + s["dir"].astype(np.float64) * s["dist"]
        b1, b2 = np.linalg.solve(G, [e1 @ (y - LIGHT[0]), e2 @ (y - LIGHT[0])])
        assert b1 > -1e-4 and b2 > -1e-4 and b1 + b2 < 1 + 1e-4 and abs(y[2] - 1.0) < 1e-4
        if i < 2000:
            p = orc.light_pdf(x, s["dir"])
            assert abs(s["pdf"] - p) <= 1e-4 * p
        # sub-triangle of a 4 x 4 subdivision (16 cells of equal area): the
        # lower cell (iu, iv), or the upper one when the fractional parts sum past 1
        k = 4
        u, v = min(max(b1, 0.0) * k, k - 1e-9), min(max(b2, 0.0) * k, k - 1e-9)
        iu, iv = int(u), int(v)
        if iu + iv > k - 1:
            iv = k - 1 - iu
        keys[(iu, iv, (u - iu) + (v - iv) > 1)] += 1
    assert len(keys) == 16, keys
    counts = np.array(list(keys.values()), np.float64)
    chi2 = float(((counts - N / 16) ** 2 / (N / 16)).sum())
    assert stats.chi2.sf(chi2, 15) > 1e-4, (chi2, counts)


def test_nee_picks_emitters_by_weight(cg, oracle):
    """With two emitters the alias pick (emission.wesl:181-191) chooses each with
    its normalised weight area * luminance: a binomial test on 40,000 picks."""
    big = np.array([[-1.0, -1.0, 1.5], [1.0, -1.0, 1.5], [0.0, 1.2, 1.5]])
    s = _light_scene(cg, second_light=big)
    orc = oracle.OracleScene.build(s.describe())
    a1, a2 = _tri_area(LIGHT), _tri_area(big)
    p1 = a1 / (a1 + a2)  # same emission, so the weights are the areas
    st = 31337
    N = 40_000
    first = 0
    for _ in range(N):
        e = orc.emission_sample(600.0, [0.1, 0.1, 0.0], st)
        st = e["state"]
        first += e["prim"] == 0
    assert stats.binomtest(first, N, p1).pvalue > 1e-4, (first / N, p1)


def test_balance_weights_sum_to_one(light_scene, oracle):
    """Balance-heuristic weights n_k p_k / sum_j n_j p_j of every sampled
    technique sum to 1 at any direction (sample.wesl:452-501): one lobe plus
    the light, and two lobes plus the light."""
    _, orc = light_scene
    rng = np.random.default_rng(17)
    x = np.array([0.2, 0.1, 0.0])
    d = unit([0.3, -0.2, -1.0])
    for counts, model_list, mat in (((1, 0, 0, 0, 0, 1), [oracle.DIFFUSE], oracle.mat(diffuse=0.6, alpha=1.0, ior=1.0)),
                                    ((1, 0, 1, 0, 0, 1), [oracle.DIFFUSE, oracle.ROUGH_REFLECT],
                                     oracle.mat(diffuse=0.6, alpha=0.4, ior=1.5))):
        for _ in range(300):
            if rng.random() < 0.5:
                w = unit(rng.dirichlet([1, 1, 1]) @ LIGHT - x)  # toward the light: p_light > 0
            else:
                w = unit(rng.normal(size=3) * [1, 1, 0.3] + [0, 0, 0.8])
            pl = orc.light_pdf(x, w)
            pk = [oracle.lobe_eval(k, d, N_UP, mat, w)[0] for k in model_list]
            den = orc.balance_denominator(d, x, N_UP, mat, counts, 0, -1.0, -1.0, w)
            if sum(pk) + pl <= 1e-5:
                continue
            weights = [p / den for p in pk] + [pl / den]
            assert abs(sum(weights) - 1.0) < 2e-6, (weights, den)
            # the sampled technique's own pdf replaces its recomputation
            den2 = orc.balance_denominator(d, x, N_UP, mat, counts, model_list[0], pk[0], -1.0, w)
            assert abs(den2 - den) <= 1e-6 * den


# --- Camera ray: compute.wesl:64-79 ---------------------------------------------


def test_camera_ray_closed_form(cg, oracle):
    """A camera looking along +x with z up: the image x axis runs toward -y
    (the camera's right), the image y axis toward -z (down), and pixel (px, py)
    looks along normalize(1, -cx vx, -cy vy) with c = 2 p / size - 1; the ray
    starts at the camera position.  Rotating the camera about z rotates every
    ray with it."""
    L = oracle.lib()
    W, H = 40, 30
    vx, vy = 0.7, 0.4
    pos = (1.5, -2.0, 0.25)
    ray = np.zeros(1, cg.RAY)
    for yaw in (0.0, 0.7, 2.5):
        cd = np.array([math.cos(yaw), math.sin(yaw), 0.0])
        cam = cg.Camera.make(pos, cd, (vx, vy))
        R = np.array([[math.cos(yaw), -math.sin(yaw), 0], [math.sin(yaw), math.cos(yaw), 0], [0, 0, 1]])
        for px, py in ((0, 0), (W // 2, H // 2), (W - 1, H - 1), (7, 23), (31, 2)):
            L.orc_camera_ray(C.byref(cam), W, H, px, py, ray.ctypes.data)
            cx, cy = 2 * px / W - 1, 2 * py / H - 1
            expect = R @ unit([1.0, -cx * vx, -cy * vy])
            assert np.abs(ray["dir"][0] - expect).max() < 2e-6, (yaw, px, py)
            assert np.all(ray["org"][0] == np.float32(pos))
    # a pitched camera: the centre pixel looks along the camera direction, the
    # image axes stay perpendicular to it, and the horizontal axis is level
    cd = unit([0.6, 0.3, -0.5])
    cam = cg.Camera.make((0, 0, 0), cd, (vx, vy))
    L.orc_camera_ray(C.byref(cam), W, H, W // 2, H // 2, ray.ctypes.data)
    assert np.abs(ray["dir"][0] - cd).max() < 2e-6
    L.orc_camera_ray(C.byref(cam), W, H, W - 1 - 4, H // 2, ray.ctypes.data)
    right = ray["dir"][0].astype(np.float64) / np.dot(ray["dir"][0], cd) - cd
    assert abs(right[2]) < 1e-6 and abs(np.dot(right, cd)) < 1e-6
    L.orc_camera_ray(C.byref(cam), W, H, W // 2, H - 1, ray.ctypes.data)
    down = ray["dir"][0].astype(np.float64) / np.dot(ray["dir"][0], cd) - cd
    assert abs(np.dot(down, cd)) < 1e-6 and abs(np.dot(down, right)) < 1e-6 and down[2] < 0


# --- Spawn: spawn_surface_ray.wesl:6-15 ----------------------------------------


def test_spawn_offsets_toward_the_next_direction(oracle):
    L = oracle.lib()
    p = [0.5, -1.0, 2.0]
    n = unit([0.2, 0.3, 0.9])
    for nxt, sign in (([0.1, 0.0, 1.0], 1.0), ([0.1, 0.0, -1.0], -1.0), (list(unit([1.0, 1.0, -0.6])), -1.0)):
        sign = 1.0 if np.dot(nxt, n) > 0 else -1.0
        o = oracle.vec_out(L.orc_spawn, oracle.fvec(p), oracle.fvec(n), oracle.fvec(nxt))
        assert np.abs(o - (np.array(p) + sign * 1e-3 * n)).max() < 1e-6


# --- Whole-path anchors -------------------------------------------------------


def _replay_wavelength(oracle, seed):
    """The pixel's first sample wavelength: seedRng steps the pixel index once,
    then one uniform goes to the wavelength sampler (render.wesl:7-18)."""
    L = oracle.lib()
    st = C.c_uint32(pcg_hash_model(seed))
    u = L.orc_rand_float(C.byref(st))
    lam, prob = C.c_float(), C.c_float()
    L.orc_sample_wavelength(u, C.byref(lam), C.byref(prob))
    return lam.value


def test_diffuse_floor_under_constant_sky_is_albedo_times_sky(cg, oracle):
    """Trace depth 2, no emitters: the camera ray hits the diffuse floor, one
    cosine-sampled bounce escapes to a constant sky, and f cos / pdf = R, so
    every sample is exactly 3 xyz(l) R(l) L(l) / sum(xyz(l)) at the pixel's
    wavelength (FP64 CMFs; R, L the reflectance and radiance lifts)."""
    L = oracle.lib()
    sky = (0.6, 0.7, 0.9)
    s = _light_scene(cg, albedo=0.8, sky=sky, with_light=False)
    orc = oracle.OracleScene.build(s.describe())
    W, H = 12, 10
    conf = cg.Conf.make(W, H, 1, 2)
    cam = cg.Camera.make((-1.5, 0.0, 1.0), (1.0, 0.1, -0.8), (0.4, 0.3))
    out = orc.render(conf, cam)
    for py in range(H):
        for px in range(W):
            lam = _replay_wavelength(oracle, px + py * W)
            R = L.orc_rgb_to_reflectance(oracle.fvec([0.8, 0.8, 0.8]), lam)
            Le = L.orc_rgb_to_radiance(oracle.fvec(sky), lam)
            xyz = np.array([cmf64(lam, LOBES[c]) for c in "xyz"])
            expect = 3 * xyz * R * Le / xyz.sum()
            got = out["history"]["xyz"][py, px]
            assert np.allclose(got, expect, rtol=2e-5, atol=1e-8), (px, py, got, expect)
            # no bounce where the Smits lift gives R = 0 (the 686-720 nm quirk)
            assert out["ray_count"][py, px] == (2 if R > 0 else 1)


def _tri_quadrature(k=96):
    """Centroids and areas (fractions of the triangle) of a k x k subdivision."""
    pts, wts = [], []
    for i in range(k):
        for j in range(k - i):
            pts.append(((i + 1 / 3) / k, (j + 1 / 3) / k))
            wts.append(1.0)
            if i + j < k - 1:
                pts.append(((i + 2 / 3) / k, (j + 2 / 3) / k))
                wts.append(1.0)
    pts = np.array(pts)
    return pts, np.array(wts) / len(wts)


def _mis_integrals(x):
    """A_T = int (cos/pi) p_d / (p_d + p_l) dw and A_N = int (cos^2/pi) p_l /
    (p_d + p_l) dw over the light seen from floor point x, with p_d = cos/pi,
    p_l = (1/A) r^2/|cos_l| and dw = |cos_l| / r^2 dA (FP64)."""
    bc, fr = _tri_quadrature()
    A = _tri_area(LIGHT)
    e1, e2 = LIGHT[1] - LIGHT[0], LIGHT[2] - LIGHT[0]
    nl = unit(np.cross(e1, e2))
    y = LIGHT[0] + bc[:, :1] * e1 + bc[:, 1:] * e2
    off = y - x
    r2 = (off * off).sum(axis=1)
    w = off / np.sqrt(r2)[:, None]
    cos = w[:, 2]
    cosl = np.abs(w @ nl)
    pd = cos / math.pi
    pl = r2 / (A * cosl)
    dw = cosl / r2 * A * fr
    at = float((cos / math.pi * pd / (pd + pl) * dw).sum())
    an = float((cos * cos / math.pi * pl / (pd + pl) * dw).sum())
    an_single_cos = float((cos / math.pi * pl / (pd + pl) * dw).sum())
    return at, an, an_single_cos


@pytest.mark.parametrize("depth", [2, 3])
def test_direct_light_mis_expectation(cg, oracle, depth):
    """One diffuse floor (IOR 1: no specular lobe), one emissive triangle, a
    black sky.  At the floor the path draws one diffuse direction and one light
    sample (counts 1 and 1) and weighs both by the balance heuristic.  Derived
    from the reference's estimators (sample.wesl:384-450, trace.wesl:22-113):
      - light sample: Le_N R (cos/pi) |n.w| / (p_d + p_l) per draw with pdf
        p_l: expectation R Le_N A_N (the reference multiplies the BSDF term,
        which already holds the cosine, by |n.w| once more);
      - diffuse sample hitting the light: R (cos/pi) / (p_d + p_l) times the
        light's radiance, which is the *reflectance* lift of its emission,
        0.94 Le_N (material.wesl:51); expectation 0.94 R Le_N A_T — counted only
        when depth 1 is not the last level (trace depth 3), else 0.
    With the wavelength independent of the geometric draws, the pixel mean is
    (3/total) int xyz R Le_N dl  x  (A_N + [depth 3] 0.94 A_T): compared per
    pixel over 192 pixels x 512 samples (5 standard errors)."""
    L = oracle.lib()
    s = _light_scene(cg, albedo=1.0)
    orc = oracle.OracleScene.build(s.describe())
    W, H, spp = 16, 12, 512
    conf = cg.Conf.make(W, H, spp, depth)
    cam = cg.Camera.make((-1.5, 0.0, 0.5), (1.0, 0.05, -0.45), (0.35, 0.25))
    out = orc.render(conf, cam, threads=8)
    # spectral factor: (3 / total) int xyz(l) R(l) Le_N(l) dl on [380, 720]
    lam = np.linspace(380.0, 720.0, 34001)
    one = oracle.fvec([1.0, 1.0, 1.0])
    R = np.array([L.orc_rgb_to_reflectance(one, float(l)) for l in lam])
    Le = np.array([L.orc_rgb_to_radiance(one, float(l)) for l in lam])
    xyz = np.array([[cmf64(float(l), LOBES[c]) for c in "xyz"] for l in lam])
    total = float(integrate.trapezoid(xyz.sum(axis=1), lam))
    K = 3 / total * integrate.trapezoid(xyz * (R * Le)[:, None], lam, axis=0)
    ratios, alt = [], []
    ray = np.zeros(1, cg.RAY)
    for py in range(H):
        for px in range(W):
            L.orc_camera_ray(C.byref(cam), W, H, px, py, ray.ctypes.data)
            h = orc.trace_closest(ray)["hits"][0]
            assert h["inst"] != MISS and h["t"] < 10
            x = ray["org"][0].astype(np.float64) + h["t"] * ray["dir"][0].astype(np.float64)
            assert abs(x[2]) < 1e-5  # the floor
            at, an, an1 = _mis_integrals(x)
            tr = 0.94 * at if depth == 3 else 0.0
            expect = K * (an + tr)
            ratios.append(out["history"]["xyz"][py, px] / expect)
            alt.append(out["history"]["xyz"][py, px] / (K * (an1 + tr)))
    r, ra = np.array(ratios), np.array(alt)
    for c in range(3):
        mean, sem = r[:, c].mean(), r[:, c].std() / math.sqrt(len(r))
        assert abs(mean - 1.0) < 5 * sem + 2e-4, (c, mean, sem)
        # the test tells the reference's estimator from the textbook one (a
        # single cosine in the light sample's term): that one is far outside
        assert abs(ra[:, c].mean() - 1.0) > 20 * sem, (c, ra[:, c].mean(), sem)
