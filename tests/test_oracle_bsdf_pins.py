"""CPU pins of the oracle's shading functions (oracle/oracle_bsdf.h) by what
the mathematics fixes, independent of the oracle's own formulas.

The functions restate PathTrace.metal: Fresnel (:352-399), GGX D and Smith G1
(:1177-1264), evaluateMaterialBRDF (:1330-1390), materialLobeProbabilities,
ggxReflectionPDF and materialBSDFPDF (:1398-1495), the diffuse bounce weight
(:1497-1523), sampleGGXHalfVector and sampleGGXVisibleNormal (:1615-1664),
the lobe sampling that ends a bounce (:5423-5513), continueTransmissiveSurface
(:1723-1799), Henyey-Greenstein (:311-324), subsurface free flight and its
weights (:3490-3523) and the transmission-volume walk (:3685-3721).

What pins each (every check is a function of the loaded pin library, so the
last test can run it on copies of the oracle with a planted mistake):
  * Fresnel: ((n-1)/(n+1))^2 at normal incidence; Fresnel's sine and tangent
    laws at every angle; total internal reflection past the critical angle.
  * GGX D: the projected-area normalisation  int D(h) (n.h) dw = 1; the peak
    D(n) = 1/(pi alpha^2) with alpha = roughness^2, independent of the
    anisotropy (the anisotropic alphas keep their product).
  * Smith G1: Heitz's closed form 1/(1 + Lambda), Lambda written with tan and
    the projected roughness (JCGT 2014, eq. 86 and sec. 5.3).
  * ggxReflectionPDF: the weak white furnace  int pdf(l) dl = 1 over the
    sphere (a D, G1 or Jacobian mistake breaks it); sampleGGXVisibleNormal
    reflected about its half vector has that pdf (chi-square histogram).
  * sampleGGXHalfVector: chi-square histogram against D(h) (n.h).
  * materialBSDFPDF: integrates to pd + ps (1 - leak_s) + pc (1 - leak_c) over
    the upper hemisphere (leaks = share of visible-normal reflections below
    the surface, counted from the sampler); the lobe sampler's directions
    follow materialBSDFPDF plus the below-surface fallback's cosine share
    (chi-square); the sampler's weights are an unbiased estimate of
    int f cos + that fallback's diffuse energy (Monte Carlo vs quadrature).
  * evaluateMaterialBRDF: reciprocity f(v, l) = f(l, v) on random layered
    materials; for specular 0, f(v, v) = albedo / pi exactly (Schlick's term
    vanishes at l = v) and the furnace at normal incidence returns the albedo;
    a white metal never returns more than it receives (G1(v) G1(l) <= G1(v)).
  * Henyey-Greenstein: mean cosine g; chi-square against the HG phase function.
  * free flight: histogram against the channel-averaged exponential (the
    collision pdf), survival = the boundary pdf, unbiased weights:
    E[w 1{d<L}] = albedo (1 - e^(-s L)) and E[w 1{d>=L}] = e^(-s L) per channel.
  * transmissive surface: Snell's law, reflection share = clamp(F, .02, .98)
    with F from the Fresnel laws, E[weight] = 1 for a white dielectric (smooth
    and rough), weight 1 under total internal reflection, Beer-Lambert
    absorption at the reference distance equals the absorption colour.
  * transmission-volume walk: E[T 1{escape}] = e^(-st t),
    E[T 1{scatter}] = ss/st (1 - e^(-st t)) per channel.
"""
import math
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
from scipy import stats

import bsdf_pins as BP
from denrimrendererkit_amd import Material

Z = BP.orthonormal_frame([0.0, 0.0, 1.0])
TILTED = BP.orthonormal_frame([0.3, -0.5, 0.8])


@pytest.fixture(scope="module")
def pins():
    return BP.Pins()


# ----------------------------------------------------------------- helpers --
def sph(theta, phi):
    st = np.sin(theta)
    return np.stack([st * np.cos(phi), st * np.sin(phi), np.cos(theta)], -1)


def sphere_grid(n_theta=1536, n_phi=768, lo=0.0, hi=math.pi):
    """Midpoint rule on theta in [lo, hi] x phi in [0, 2 pi): directions (in
    the frame's z-up coordinates) and solid-angle weights."""
    dt, dp = (hi - lo) / n_theta, 2 * math.pi / n_phi
    th = lo + (np.arange(n_theta) + 0.5) * dt
    ph = (np.arange(n_phi) + 0.5) * dp
    T, P = np.meshgrid(th, ph, indexing="ij")
    return sph(T, P).reshape(-1, 3), (np.sin(T) * dt * dp).reshape(-1)


def to_world(frame, local):
    n, t, b = frame[:3], frame[3:6], frame[6:]
    local = np.asarray(local, np.float64)
    return local[..., :1] * t + local[..., 1:2] * b + local[..., 2:3] * n


def to_local(frame, world):
    n, t, b = frame[:3], frame[3:6], frame[6:]
    return np.stack([world @ t, world @ b, world @ n], -1)


def view_dir(theta_deg, phi_deg=0.0, frame=Z):
    return to_world(frame, sph(math.radians(theta_deg), math.radians(phi_deg)))


def reflect_about(v, h):
    return 2.0 * (h @ v)[:, None] * h - v[None, :]


def uniforms(n, seed, k=2):
    return np.random.default_rng(seed).random((n, k))


def chi2_pvalue(observed, expected, min_expected=5.0):
    """Pearson chi-square over bins with expected >= min_expected, the rest
    pooled into one bin."""
    observed, expected = np.asarray(observed, float).ravel(), np.asarray(expected, float).ravel()
    keep = expected >= min_expected
    o = np.append(observed[keep], observed[~keep].sum())
    e = np.append(expected[keep], expected[~keep].sum())
    if e[-1] < min_expected:
        o, e = o[:-1], e[:-1]
    e = e * (o.sum() / e.sum())
    chi = float(((o - e) ** 2 / e).sum())
    return float(stats.chi2.sf(chi, len(o) - 1))


def bin_expected(pdf_fn, frame, n_theta, n_phi, sub=12, theta_max=math.pi):
    """N-independent expected mass of each (theta, phi) bin of a density over
    directions, by a sub x sub midpoint rule inside each bin."""
    dt, dp = theta_max / n_theta, 2 * math.pi / n_phi
    th = (np.arange(n_theta * sub) + 0.5) * dt / sub
    ph = (np.arange(n_phi * sub) + 0.5) * dp / sub
    T, P = np.meshgrid(th, ph, indexing="ij")
    w = np.sin(T) * (dt / sub) * (dp / sub)
    val = pdf_fn(to_world(frame, sph(T, P).reshape(-1, 3))).reshape(T.shape) * w
    return val.reshape(n_theta, sub, n_phi, sub).sum(axis=(1, 3))


def bin_counts(dirs, frame, n_theta, n_phi, theta_max=math.pi):
    loc = to_local(frame, dirs)
    th = np.arccos(np.clip(loc[:, 2], -1, 1))
    ph = np.mod(np.arctan2(loc[:, 1], loc[:, 0]), 2 * math.pi)
    it = np.minimum((th / (theta_max / n_theta)).astype(int), n_theta - 1)
    ip = np.minimum((ph / (2 * math.pi / n_phi)).astype(int), n_phi - 1)
    counts = np.zeros((n_theta, n_phi))
    np.add.at(counts, (it, ip), 1)
    return counts


def mc_mean_ok(samples, expected, sigmas=5.0, floor=1e-6):
    """Monte Carlo mean within `sigmas` standard errors (per column)."""
    samples = np.asarray(samples, float)
    mean = samples.mean(0)
    se = samples.std(0) / math.sqrt(len(samples))
    return np.all(np.abs(mean - expected) <= sigmas * se + floor), mean, se


def mat(**kw):
    return Material(**kw).gpu_material()


def fresnel_laws(cos_i, eta_i, eta_t):
    """Fresnel's sine and tangent laws (unpolarised), independent of the
    oracle's (n cos) form; 1 under total internal reflection."""
    cos_i = np.clip(cos_i, 0.0, 1.0)
    ti = np.arccos(cos_i)
    s = eta_i / eta_t * np.sin(ti)
    out = np.ones_like(cos_i)
    ok = s < 1.0
    tt = np.arcsin(s[ok])
    a, b = ti[ok] - tt, ti[ok] + tt
    with np.errstate(divide="ignore", invalid="ignore"):
        rs = np.where(b > 0, -np.sin(a) / np.sin(b), (eta_i - eta_t) / (eta_i + eta_t))
        rp = np.where(b > 0, np.tan(a) / np.tan(b), (eta_t - eta_i) / (eta_t + eta_i))
    out[ok] = 0.5 * (rs * rs + rp * rp)
    return out


# ------------------------------------------------------------------ checks --
def check_fresnel(P):
    for n in (1.1, 1.33, 1.5, 1.7, 2.4):
        f0 = ((n - 1) / (n + 1)) ** 2
        assert abs(P.dielectric_fresnel([1.0], 1.0, n)[0] - f0) < 1e-12, "normal incidence, outside"
        assert abs(P.dielectric_fresnel([1.0], n, 1.0)[0] - f0) < 1e-12, "normal incidence, inside"
    c = np.linspace(0.002, 0.999, 400)
    for ei, et in ((1.0, 1.5), (1.5, 1.0), (1.0, 1.33), (2.0, 1.2)):
        got = P.dielectric_fresnel(c, ei, et)
        assert np.allclose(got, fresnel_laws(c, ei, et), rtol=1e-9, atol=1e-12), (ei, et)
    crit = math.sqrt(1 - (1 / 1.5) ** 2)  # cos of the critical angle from glass
    assert np.all(P.dielectric_fresnel(np.linspace(0.0, crit - 1e-6, 50), 1.5, 1.0) == 1.0), "TIR"
    assert P.dielectric_fresnel([crit + 1e-3], 1.5, 1.0)[0] < 1.0


def check_ggx_ndf(P):
    dirs, w = sphere_grid(1536, 512, 0.0, math.pi / 2)
    for r in (0.3, 0.5, 0.8):
        a = r * r
        iso = P.ggx_d(Z, r, 0.0, dirs, anisotropic=False)
        assert abs((iso * dirs[:, 2] * w).sum() - 1) < 2e-3, f"isotropic D not normalised at r={r}"
        assert abs(P.ggx_d(Z, r, 0.0, [[0, 0, 1]], anisotropic=False)[0] * math.pi * a * a - 1) < 1e-9
        for an in (0.0, 0.7, -0.7):
            d = P.ggx_d(TILTED, r, an, to_world(TILTED, dirs))
            assert abs((d * dirs[:, 2] * w).sum() - 1) < 2e-3, f"anisotropic D not normalised at r={r}, {an}"
            peak = P.ggx_d(TILTED, r, an, TILTED[None, :3])[0]
            assert abs(peak * math.pi * a * a - 1) < 1e-9, "D(n) = 1/(pi alpha^2) for any anisotropy"
        # positive anisotropy widens the lobe along the tangent, negative along the bitangent
        ht = to_world(TILTED, sph(0.25, 0.0)[None])
        hb = to_world(TILTED, sph(0.25, math.pi / 2)[None])
        assert P.ggx_d(TILTED, r, 0.7, ht)[0] > P.ggx_d(TILTED, r, 0.7, hb)[0]
        assert P.ggx_d(TILTED, r, -0.7, ht)[0] < P.ggx_d(TILTED, r, -0.7, hb)[0]


def burley_alphas(r, an):
    """Burley 2012 anisotropic roughness (aspect = sqrt(1 - 0.9 |a|)), floored at 1e-3."""
    a = max(r * r, 1e-3)
    aspect = math.sqrt(max(0.1, 1 - 0.9 * abs(an)))
    return (a / aspect, a * aspect) if an >= 0 else (a * aspect, a / aspect)


def smith_g1_closed_form(local, ax, ay):
    cos_t = local[:, 2]
    tan2 = np.maximum(1 - cos_t ** 2, 0) / cos_t ** 2
    phi = np.arctan2(local[:, 1], local[:, 0])
    a2 = np.cos(phi) ** 2 * ax * ax + np.sin(phi) ** 2 * ay * ay
    lam = (-1 + np.sqrt(1 + a2 * tan2)) / 2
    return 1 / (1 + lam)


def check_smith_g1(P):
    rng = np.random.default_rng(3)
    th = np.arccos(rng.uniform(0.02, 1.0, 2000))
    ph = rng.uniform(0, 2 * math.pi, 2000)
    local = sph(th, ph)
    for r in (0.2, 0.5, 0.9):
        got = P.ggx_g1(Z, r, 0.0, local, anisotropic=False)
        assert np.allclose(got, smith_g1_closed_form(local, r * r, r * r), rtol=1e-9), f"isotropic G1 r={r}"
        for an in (0.0, 0.7, -0.5):
            ax, ay = burley_alphas(r, an)
            got = P.ggx_g1(TILTED, r, an, to_world(TILTED, local))
            assert np.allclose(got, smith_g1_closed_form(local, ax, ay), rtol=1e-9), f"anisotropic G1 r={r} a={an}"
    assert abs(P.ggx_g1(Z, 0.5, 0.0, [[0, 0, 1]], anisotropic=False)[0] - 1) < 1e-12


VNDF_CASES = [(0.3, 0.0, 0.0, 0.0), (0.3, 0.0, 60.0, 0.0), (0.6, 0.7, 45.0, 30.0), (0.45, -0.7, 70.0, 120.0),
              (0.8, 0.7, 20.0, 200.0)]


def check_ggx_pdf_normalised(P):
    dirs, w = sphere_grid(1536, 768)
    for r, an, tv, pv in VNDF_CASES:
        v = view_dir(tv, pv, TILTED)
        p = P.ggx_pdf(TILTED, r, an, np.broadcast_to(v, dirs.shape), to_world(TILTED, dirs))
        total = (p * w).sum()
        assert abs(total - 1) < 3e-3, f"int ggxReflectionPDF = {total} (r={r} a={an} view {tv},{pv})"


def check_vndf_sampling(P):
    n, nt, npf = 300_000, 36, 36
    for k, (r, an, tv, pv) in enumerate(VNDF_CASES):
        v = view_dir(tv, pv, TILTED)
        h = P.sample_vndf(TILTED, r, an, v, uniforms(n, 10 + k))
        l = reflect_about(v, h)
        exp = bin_expected(lambda d: P.ggx_pdf(TILTED, r, an, np.broadcast_to(v, d.shape), d), TILTED, nt, npf)
        assert abs(exp.sum() - 1) < 5e-3
        p = chi2_pvalue(bin_counts(l, TILTED, nt, npf), exp * n)
        assert p > 1e-4, f"visible-normal sampling vs ggxReflectionPDF: p={p:.2e} (r={r} a={an} view {tv},{pv})"


def check_ggx_half_vector_sampling(P):
    n, nt, npf = 300_000, 40, 16
    for k, r in enumerate((0.35, 0.7)):
        h = P.sample_ggx_half_vector(TILTED[:3], r, uniforms(n, 30 + k))
        nh = lambda d: P.ggx_d(TILTED, r, 0.0, d, anisotropic=False) * np.clip(d @ TILTED[:3], 0, None)
        exp = bin_expected(nh, TILTED, nt, npf, theta_max=math.pi / 2)
        assert abs(exp.sum() - 1) < 5e-3
        p = chi2_pvalue(bin_counts(h, TILTED, nt, npf, theta_max=math.pi / 2), exp * n)
        assert p > 1e-4, f"GGX half-vector sampling vs D(h)(n.h): p={p:.2e} (r={r})"


LAYERED = dict(base_color=(0.6, 0.35, 0.2), roughness=0.45, specular=0.8, specular_anisotropy=0.6, clearcoat=0.7,
               clearcoat_roughness=0.3, sheen=0.5, sheen_color=(0.9, 0.8, 0.7), sheen_roughness=0.4,
               thin_film=0.6, thin_film_thickness_nanometers=420.0, clearcoat_thickness=0.2,
               clearcoat_attenuation_color=(0.8, 0.6, 0.5))
PDF_MATERIALS = {
    "diffuse": dict(base_color=(0.7, 0.7, 0.7), specular=0.0),
    "metal": dict(base_color=(0.9, 0.9, 0.9), metallic=1.0, roughness=0.55),
    "metal_aniso": dict(base_color=(0.9, 0.9, 0.9), metallic=1.0, roughness=0.55, specular_anisotropy=0.7),
    "clearcoat": dict(base_color=(0.0, 0.0, 0.0), specular=0.0, clearcoat=1.0, clearcoat_roughness=0.4),
    "coated": dict(base_color=(0.5, 0.3, 0.2), specular=0.6, roughness=0.35, clearcoat=0.8, clearcoat_roughness=0.25),
    "layered": LAYERED,
}


def vndf_leak(P, r, an, v, n=400_000, seed=50):
    """Share of visible-normal reflections below the surface, counted from the sampler."""
    h = P.sample_vndf(TILTED, r, an, v, uniforms(n, seed))
    return float((reflect_about(v, h) @ TILTED[:3] <= 0).mean()), n


def material_lobes(P, kw):
    m = mat(**kw)
    r = min(max(kw.get("roughness", 0.5), 0.02), 1.0)
    an = max(min(kw.get("specular_anisotropy", 0.0), 0.95), -0.95)
    rc = min(max(kw.get("clearcoat_roughness", 0.1), 0.02), 1.0)
    return m, P.lobe_probabilities(m), (r, an), (rc, 0.0)


def check_material_pdf_normalisation(P):
    up, w = sphere_grid(1024, 768, 0.0, math.pi / 2)
    lw = to_world(TILTED, up)
    for name, kw in PDF_MATERIALS.items():
        for tv in (0.0, 55.0):
            v = view_dir(tv, 25.0, TILTED)
            m, (pd, ps, pc), spec, coat = material_lobes(P, kw)
            assert abs(pd + ps + pc - 1) < 1e-12 and min(pd, ps, pc) >= 0
            total = (P.pdf(m, TILTED, np.broadcast_to(v, lw.shape), lw) * w).sum()
            ls, ns = vndf_leak(P, *spec, v) if ps > 0 else (0.0, 1)
            lc, nc = vndf_leak(P, *coat, v, seed=51) if pc > 0 else (0.0, 1)
            expect = pd + ps * (1 - ls) + pc * (1 - lc)
            tol = 2e-3 + 5 * (ps * math.sqrt(ls * (1 - ls) / ns) + pc * math.sqrt(lc * (1 - lc) / nc))
            assert abs(total - expect) < tol, f"{name} view {tv}: int pdf {total:.5f} vs {expect:.5f}"


def _fallback_share(P, m, spec, coat, lobes, v):
    pd, ps, pc = lobes
    dirs, w = sphere_grid(1536, 768, math.pi / 2, math.pi)  # lower hemisphere
    lw = to_world(TILTED, dirs)
    vv = np.broadcast_to(v, lw.shape)
    ls = (P.ggx_pdf(TILTED, *spec, vv, lw) * w).sum() if ps > 0 else 0.0
    lc = (P.ggx_pdf(TILTED, *coat, vv, lw) * w).sum() if pc > 0 else 0.0
    return ps * ls + pc * lc


def check_bounce_sampler_density(P):
    n, nt, npf = 400_000, 24, 32
    for k, (name, tv) in enumerate((("layered", 35.0), ("metal_aniso", 60.0), ("clearcoat", 20.0))):
        m, lobes, spec, coat = material_lobes(P, PDF_MATERIALS[name])
        v = view_dir(tv, 40.0, TILTED)
        d, _ = P.sample_bounce(m, TILTED, -v, n, seed=100 + k)
        assert np.all(d @ TILTED[:3] > 0), "the bounce sampler leaves above the surface"
        fb = _fallback_share(P, m, spec, coat, lobes, v)
        dens = lambda l: (P.pdf(m, TILTED, np.broadcast_to(v, l.shape), l) +
                          fb * np.clip(l @ TILTED[:3], 0, None) / math.pi)
        exp = bin_expected(dens, TILTED, nt, npf, theta_max=math.pi / 2)
        assert abs(exp.sum() - 1) < 5e-3, f"{name}: density mass {exp.sum()}"
        p = chi2_pvalue(bin_counts(d, TILTED, nt, npf, theta_max=math.pi / 2), exp * n)
        assert p > 1e-4, f"{name}: bounce directions vs materialBSDFPDF + fallback: p={p:.2e}"


def check_bounce_sampler_estimator(P):
    """E[weight] of the lobe sampler = int f cos over the upper hemisphere
    plus the below-surface fallback's diffuse energy, for materials whose
    every nonzero BRDF term has a lobe that samples it. With specular 0 the
    specular lobe has probability 0 while the evaluation keeps Schlick's
    (1 - v.h)^5 term, so the reference's estimate is the diffuse part alone
    (the diffuse bounce weight / pi); the "diffuse" case checks that."""
    up, w = sphere_grid(1536, 768, 0.0, math.pi / 2)
    lw = to_world(TILTED, up)
    n = 1_000_000
    cases = (("layered", 35.0), ("metal", 50.0), ("metal_aniso", 60.0), ("coated", 10.0), ("diffuse", 70.0))
    for k, (name, tv) in enumerate(cases):
        m, lobes, spec, coat = material_lobes(P, PDF_MATERIALS[name])
        v = view_dir(tv, 40.0, TILTED)
        vv = np.broadcast_to(v, lw.shape)
        cos = np.clip(lw @ TILTED[:3], 0, None)
        dbw = P.diffuse_bounce_weight(m, TILTED, vv, lw)
        diffuse_energy = (dbw * (cos * w / math.pi)[:, None]).sum(0)
        if lobes[1] == 0 and lobes[2] == 0:
            energy = diffuse_energy
        else:
            energy = (P.brdf(m, TILTED, vv, lw) * (cos * w)[:, None]).sum(0)
            # the fallback divides by the diffuse probability, floored at 1e-5 as the sampler does
            energy = energy + _fallback_share(P, m, spec, coat, lobes, v) / max(lobes[0], 1e-5) * diffuse_energy
        _, wts = P.sample_bounce(m, TILTED, -v, n, seed=200 + k)
        ok, mean, se = mc_mean_ok(wts, energy, floor=2e-3 * energy.max())
        assert ok, f"{name}: E[weight] {mean} vs int f cos (+ fallback) {energy} (se {se})"


def check_reciprocity(P):
    rng = np.random.default_rng(7)
    for k in range(24):
        kw = dict(base_color=tuple(rng.uniform(0, 1, 3)), roughness=rng.uniform(0.05, 1), metallic=rng.uniform(0, 1),
                  specular=rng.uniform(0, 1), specular_anisotropy=rng.uniform(-0.9, 0.9),
                  clearcoat=rng.uniform(0, 1), clearcoat_roughness=rng.uniform(0.05, 0.8),
                  clearcoat_thickness=rng.uniform(0, 0.5), clearcoat_attenuation_color=tuple(rng.uniform(0.2, 1, 3)),
                  sheen=rng.uniform(0, 1), sheen_roughness=rng.uniform(0, 1), thin_film=rng.uniform(0, 1),
                  thin_film_thickness_nanometers=rng.uniform(200, 800), index_of_refraction=rng.uniform(1.2, 2))
        m = mat(**kw)
        fr = BP.orthonormal_frame(rng.normal(size=3))
        a = to_world(fr, sph(np.arccos(rng.uniform(0.05, 1, 200)), rng.uniform(0, 2 * math.pi, 200)))
        b = to_world(fr, sph(np.arccos(rng.uniform(0.05, 1, 200)), rng.uniform(0, 2 * math.pi, 200)))
        for prec, tol in ((64, 1e-10), (32, 2e-5)):
            f1, f2 = P.brdf(m, fr, a, b, prec), P.brdf(m, fr, b, a, prec)
            assert np.allclose(f1, f2, rtol=tol, atol=tol * 1e-3), f"BRDF not reciprocal (material {k}, fp{prec})"


def check_white_furnace(P):
    """With specular 0 the only specular term left is Schlick's (1 - v.h)^5,
    which vanishes at l = v: there f = albedo / pi exactly. At normal
    incidence that term moves < 1e-4 of the energy, so the furnace returns
    the albedo. A white metal (F = 1) returns at most what it receives,
    because G1(v) G1(l) <= G1(v) and int D G1(v) / (4 n.v) dl = 1."""
    up, w = sphere_grid(1024, 512, 0.0, math.pi / 2)
    lw = to_world(TILTED, up)
    for albedo in ((0.7, 0.7, 0.7), (0.2, 0.5, 0.9), (1.0, 1.0, 1.0)):
        m = mat(base_color=albedo, specular=0.0)
        for tv in (0.0, 45.0, 80.0):
            v = view_dir(tv, 10.0, TILTED)
            for prec, tol in ((64, 1e-10), (32, 1e-6)):
                f = P.brdf(m, TILTED, v[None], v[None], prec)[0]
                assert np.allclose(f, np.array(albedo) / math.pi, rtol=tol), "f(v, v) = albedo / pi"
        v = TILTED[:3]
        e = (P.brdf(m, TILTED, np.broadcast_to(v, lw.shape), lw) * ((lw @ TILTED[:3]) * w)[:, None]).sum(0)
        assert np.allclose(e, albedo, rtol=1e-3), f"white furnace at normal incidence {e} vs albedo {albedo}"
    for r in (0.1, 0.4, 0.9):
        m = mat(base_color=(1, 1, 1), metallic=1.0, roughness=r)
        for tv in (0.0, 60.0, 85.0):
            v = view_dir(tv, 0.0, TILTED)
            e = (P.brdf(m, TILTED, np.broadcast_to(v, lw.shape), lw) * ((lw @ TILTED[:3]) * w)[:, None]).sum(0)
            assert np.all(e <= 1.0 + 3e-3), f"white metal reflects {e} at r={r}, view {tv}"


def hg_density(g, cos_t):
    return (1 - g * g) / (2 * (1 + g * g - 2 * g * cos_t) ** 1.5)


def check_henyey_greenstein(P):
    n = 400_000
    axis = np.array([0.48, -0.6, 0.64])
    axis = axis / np.linalg.norm(axis)
    for k, g in enumerate((-0.8, -0.3, 0.0, 0.3, 0.8)):
        d = P.sample_hg(g, axis, uniforms(n, 60 + k))
        assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-12)
        c = d @ axis
        ok, mean, se = mc_mean_ok(c[:, None], np.array([g]))
        assert ok, f"HG mean cosine {mean} vs g={g}"
        edges = np.linspace(-1, 1, 41)
        sub = np.linspace(-1, 1, 40 * 64 + 1)
        mid = 0.5 * (sub[1:] + sub[:-1])
        exp = (hg_density(g, mid) * (sub[1] - sub[0])).reshape(40, 64).sum(1)
        p = chi2_pvalue(np.histogram(c, edges)[0], exp * n)
        assert p > 1e-4, f"HG cosines vs the phase function: p={p:.2e} (g={g})"


SIGMA = np.array([2.0, 5.0, 11.0])


def check_free_flight(P):
    n = 600_000
    d = P.free_flight(SIGMA, uniforms(n, 70))
    ok, mean, se = mc_mean_ok(d[:, None], np.array([np.mean(1 / SIGMA)]))
    assert ok, f"free-flight mean {mean} vs mean of 1/sigma {np.mean(1 / SIGMA)}"
    edges = np.linspace(0, 2.0, 81)
    sub = np.linspace(0, 2.0, 80 * 32 + 1)
    coll, _ = P.sss_pdfs(SIGMA, 0.5 * (sub[1:] + sub[:-1]))
    exp = (coll * (sub[1] - sub[0])).reshape(80, 32).sum(1)
    exp = np.append(exp, 1 - exp.sum())
    obs = np.append(np.histogram(d, edges)[0], (d >= 2.0).sum())
    p = chi2_pvalue(obs, exp * n)
    assert p > 1e-4, f"free flight vs the collision pdf: p={p:.2e}"
    for L in (0.05, 0.3, 1.0):
        _, bound = P.sss_pdfs(SIGMA, [L])
        ok, share, se = mc_mean_ok((d >= L)[:, None].astype(float), bound)
        assert ok, f"P(d >= {L}) = {share} vs boundary pdf {bound}"


def check_subsurface_weights(P):
    n = 600_000
    albedo = np.array([0.9, 0.6, 0.3])
    d = P.free_flight(SIGMA, uniforms(n, 80))
    for L in (0.05, 0.3, 1.0):
        coll, bound = P.sss_weights(SIGMA, albedo, np.minimum(d, L))
        inside = (d < L)[:, None]
        ok, mean, se = mc_mean_ok(np.where(inside, coll, 0.0), albedo * (1 - np.exp(-SIGMA * L)))
        assert ok, f"E[collision weight, d<{L}] {mean}"
        ok, mean, se = mc_mean_ok(np.where(inside, 0.0, bound), np.exp(-SIGMA * L))
        assert ok, f"E[boundary weight, d>={L}] {mean}"


def check_transmissive_surface(P):
    n = 200_000
    nrm = np.array([0.0, 0.0, 1.0])
    for ti in (0.0, 30.0, 60.0, 80.0):
        inc = -view_dir(ti, 0.0)
        m = mat(base_color=(1, 1, 1), transmission=1.0, transmission_color=(1, 1, 1), roughness=0.005,
                index_of_refraction=1.5)
        d, w, taken = P.transmissive_surface(m, nrm, True, 1.0, inc, n, seed=90)
        assert taken.all()
        refr = d[:, 2] < 0
        st = np.hypot(d[refr, 0], d[refr, 1])
        assert np.allclose(st, math.sin(math.radians(ti)) / 1.5, atol=1e-12), "Snell's law"
        assert np.allclose(d[refr, 1], 0, atol=1e-12) and np.all(d[refr, 0] * inc[0] >= 0), "plane of incidence"
        assert np.allclose(d[~refr], inc * [1, 1, -1], atol=1e-12), "mirror reflection"
        f = fresnel_laws(np.array([math.cos(math.radians(ti))]), 1.0, 1.5)[0]
        pr = min(max(f, 0.02), 0.98)
        ok, share, se = mc_mean_ok((~refr)[:, None].astype(float), np.array([pr]))
        assert ok, f"reflection share {share} vs clamp(F) {pr} at {ti} deg"
        assert np.allclose(w[~refr], f / pr) and np.allclose(w[refr], (1 - f) / (1 - pr)), "branch weights"
        # rough dielectric: the microfacet choice keeps E[weight] = 1 for a white material
        mr = mat(base_color=(1, 1, 1), transmission=1.0, transmission_color=(1, 1, 1), roughness=0.5,
                 index_of_refraction=1.5)
        _, wr, _ = P.transmissive_surface(mr, nrm, True, 1.0, inc, n, seed=91)
        ok, mean, se = mc_mean_ok(wr, np.ones(3))
        assert ok, f"rough dielectric E[weight] {mean} at {ti} deg"
    # from inside the glass past the critical angle: always reflected, weight 1
    m = mat(base_color=(1, 1, 1), transmission=1.0, transmission_color=(1, 1, 1), roughness=0.005,
            index_of_refraction=1.5)
    inc = -view_dir(50.0, 0.0)
    d, w, _ = P.transmissive_surface(m, nrm, False, 1.0, inc, 5000, seed=92)
    assert np.all(d[:, 2] > 0) and np.all(w == 1.0), "total internal reflection"
    # Beer-Lambert on the way out: at the reference distance the absorption equals the colour
    col, dist = np.array([0.5, 0.25, 0.8]), 0.7
    ma = mat(base_color=(1, 1, 1), transmission=1.0, transmission_color=(1, 1, 1), roughness=0.005,
             index_of_refraction=1.5, transmission_absorption_color=tuple(col), transmission_absorption_distance=dist)
    inc = -view_dir(20.0, 0.0)
    d, w, _ = P.transmissive_surface(ma, nrm, False, dist, inc, 5000, seed=93)
    out = d[:, 2] < 0  # refracted out of the glass (the normal faces the incoming side)
    assert out.any()
    f = fresnel_laws(np.array([math.cos(math.radians(20.0))]), 1.5, 1.0)[0]
    pr = min(max(f, 0.02), 0.98)
    assert np.allclose(w[out] / ((1 - f) / (1 - pr)), col, rtol=1e-9), "absorption at the reference distance"


def check_transmission_volume_walk(P):
    n = 600_000
    kw = dict(base_color=(1, 1, 1), transmission=1.0, roughness=0.005, volume_scattering=0.6,
              volume_scattering_color=(0.9, 0.7, 0.5), volume_scattering_distance=0.4,
              transmission_absorption_color=(0.6, 0.8, 0.9), transmission_absorption_distance=1.0)
    m = mat(**kw)
    ss = np.array(kw["volume_scattering_color"]) * kw["volume_scattering"] / kw["volume_scattering_distance"]
    sa = -np.log(np.array(kw["transmission_absorption_color"])) / kw["transmission_absorption_distance"]
    st = sa + ss
    for t in (0.1, 0.5, 2.0):
        w, sc = P.transmission_volume_walk(m, t, [0.0, 0.0, 1.0], n, seed=95)
        esc = (sc == 0)[:, None]
        ok, mean, se = mc_mean_ok(np.where(esc, w, 0.0), np.exp(-st * t))
        assert ok, f"E[T, escaped] {mean} vs exp(-sigma_t t) at t={t}"
        ok, mean, se = mc_mean_ok(np.where(esc, 0.0, w), ss / st * (1 - np.exp(-st * t)))
        assert ok, f"E[T, scattered] {mean} at t={t}"


def check_precisions_agree(P):
    rng = np.random.default_rng(11)
    m = mat(**LAYERED)
    a = to_world(TILTED, sph(np.arccos(rng.uniform(0.1, 1, 500)), rng.uniform(0, 2 * math.pi, 500)))
    b = to_world(TILTED, sph(np.arccos(rng.uniform(0.1, 1, 500)), rng.uniform(0, 2 * math.pi, 500)))
    for fn in (P.brdf, P.pdf):
        f64, f32 = fn(m, TILTED, a, b, 64), fn(m, TILTED, a, b, 32)
        assert np.allclose(f32, f64, rtol=2e-4, atol=1e-6), f"{fn.__name__}: fp32 and fp64 oracles disagree"


CHECKS = {f.__name__: f for f in (
    check_fresnel, check_ggx_ndf, check_smith_g1, check_ggx_pdf_normalised, check_vndf_sampling,
    check_ggx_half_vector_sampling, check_material_pdf_normalisation, check_bounce_sampler_density,
    check_bounce_sampler_estimator, check_reciprocity, check_white_furnace, check_henyey_greenstein,
    check_free_flight, check_subsurface_weights, check_transmissive_surface, check_transmission_volume_walk,
    check_precisions_agree)}


@pytest.mark.parametrize("name", list(CHECKS))
def test_oracle_pin(pins, name):
    CHECKS[name](pins)


# ---------------------------------------------------- planted mistakes --
# Each entry plants one plausible slip in a copy of oracle_bsdf.h (the text to
# replace must occur exactly once) and names the checks that must catch it.
MUTANTS = {
    "alpha_not_squared": ("R alpha = roughness * roughness;\n    R alpha2 = alpha * alpha;\n    R denominator = nDotH",
                          "R alpha = roughness;\n    R alpha2 = alpha * alpha;\n    R denominator = nDotH",
                          ["check_ggx_ndf"]),
    "pdf_missing_4": ("/ fmx(R(4) * nDotV, R(1e-5));", "/ fmx(nDotV, R(1e-5));",
                      ["check_ggx_pdf_normalised", "check_vndf_sampling"]),
    "fresnel_eta_swapped": ("    R eta = etaI / etaT;\n    R sinThetaTSquared", "    R eta = etaT / etaI;\n    R sinThetaTSquared",
                            ["check_fresnel"]),
    "diffuse_missing_inv_pi": ("(R(1) - metallic) * baseColor * inversePi;", "(R(1) - metallic) * baseColor;",
                               ["check_white_furnace"]),
    "g1_axes_swapped": ("alpha.x * alpha.x * tDotD * tDotD + alpha.y * alpha.y * bDotD * bDotD",
                        "alpha.x * alpha.x * bDotD * bDotD + alpha.y * alpha.y * tDotD * tDotD",
                        ["check_smith_g1", "check_ggx_pdf_normalised"]),
    "vndf_stretch_swapped": ("normalize(V3<R>(alpha.x * localView.x, alpha.y * localView.y, localView.z))",
                             "normalize(V3<R>(alpha.y * localView.x, alpha.x * localView.y, localView.z))",
                             ["check_vndf_sampling"]),
    "hg_sign_flipped": ("(R(1) + g * g - q * q) / (R(2) * g)", "(R(1) + g * g - q * q) / (R(-2) * g)",
                        ["check_henyey_greenstein"]),
    "collision_pdf_missing_third": ("sigmaT.z * tr.z) * R(0.33333333333)", "sigmaT.z * tr.z)",
                                    ["check_free_flight", "check_subsurface_weights"]),
    "boundary_weight_unweighted": ("return tr / subsurfaceBoundaryPDF(sigmaT, d);", "return tr;",
                                   ["check_subsurface_weights"]),
    "diffuse_pdf_uniform": ("R diffusePDF = nDotL * inversePi;", "R diffusePDF = inversePi;",
                            ["check_material_pdf_normalisation"]),
    "specular_weight_by_pd": ("fmx(ps, R(1e-5));\n        }", "fmx(pd, R(1e-5));\n        }",
                              ["check_bounce_sampler_estimator"]),
    "eval_fresnel_by_ndotv": ("V3<R> fresnel = fresnelSchlickThinFilm(vDotH, f0, m);\n    Frame<R> frame",
                              "V3<R> fresnel = fresnelSchlickThinFilm(nDotV, f0, m);\n    Frame<R> frame",
                              ["check_reciprocity"]),
    "refraction_eta_inverted": ("R eta = etaI / etaT;\n    R cosThetaI = fmx(dot(view, micro), R(0));",
                                "R eta = etaT / etaI;\n    R cosThetaI = fmx(dot(view, micro), R(0));",
                                ["check_transmissive_surface"]),
    "reflect_weight_unnormalised": ("T *= tir ? R(1) : fresnel / fmx(reflectProbability, R(1e-5));",
                                    "T *= tir ? R(1) : fresnel;", ["check_transmissive_surface"]),
    "volume_escape_unweighted": ("T *= materialVolumeTransmittance(m, b.t) / subsurfaceBoundaryPDF(sigmaT, b.t);",
                                 "T *= materialVolumeTransmittance(m, b.t);", ["check_transmission_volume_walk"]),
    "sheen_not_reciprocal": ("R(1) / fmx(R(4) * (nDotL + nDotV - nDotL * nDotV), R(1e-5));",
                             "R(1) / fmx(R(4) * (nDotL + nDotV - nDotV * nDotV), R(1e-5));", ["check_reciprocity"]),
}


def test_pins_catch_planted_mistakes(tmp_path):
    src = open(os.path.join(BP.HERE, "oracle_bsdf.h")).read()
    jobs = []
    for name, (old, new, _) in MUTANTS.items():
        assert src.count(old) == 1, f"mutant {name}: the text to replace must occur once"
        d = tmp_path / name
        d.mkdir()
        (d / "oracle_bsdf.h").write_text(src.replace(old, new))
        shutil.copy(os.path.join(BP.HERE, "oracle_pins.cpp"), d / "oracle_pins.cpp")
        jobs.append((name, str(d), str(d / f"liboracle_pins_{name}.so")))
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        list(ex.map(lambda j: BP.build(j[1], j[2]), jobs))
    survivors = []
    for name, _, lib in jobs:
        P = BP.Pins(lib)
        caught = []
        for check in MUTANTS[name][2]:
            try:
                CHECKS[check](P)
            except AssertionError:
                caught.append(check)
        if not caught:
            survivors.append(name)
    assert not survivors, f"planted mistakes no pin caught: {survivors}"
