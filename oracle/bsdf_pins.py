"""ctypes wrapper of liboracle_pins.so: the oracle's BSDF, phase-function and
free-flight functions (oracle_bsdf.h) behind C entry points (oracle_pins.cpp).

TEST INFRASTRUCTURE ONLY: tests/test_oracle_bsdf_pins.py loads it to check
those functions against what the mathematics fixes. `build(src_dir, out)`
builds the library from the sources in `src_dir` (by default this folder), so
a test can build a copy with a planted mistake into a temporary directory.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_pins.so")

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int32)


def build(src_dir: str = HERE, out: str = LIB_PATH, force: bool = False) -> str:
    flags = ["-B"] if force or src_dir != HERE else []
    subprocess.run(["make", "-s", *flags, "-C", HERE, f"PINS_SRC={src_dir}", f"PINS_OUT={out}", out], check=True)
    return out


def _p(a: np.ndarray):
    return a.ctypes.data_as(_D)


def _f64(a, shape=None) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    return a if shape is None else a.reshape(shape)


def orthonormal_frame(n) -> np.ndarray:
    """(normal, tangent, bitangent) as 9 doubles, right-handed."""
    n = np.asarray(n, np.float64)
    n = n / np.linalg.norm(n)
    helper = np.array([0.0, 1.0, 0.0]) if abs(n[1]) < 0.9 else np.array([1.0, 0.0, 0.0])
    t = np.cross(helper, n)
    t /= np.linalg.norm(t)
    b = np.cross(n, t)
    return np.concatenate([n, t, b])


class Pins:
    """One loaded pin library. Every method takes `prec` (32 or 64)."""

    def __init__(self, path: str = LIB_PATH):
        if path == LIB_PATH:
            build()
        self.path = path
        L = C.CDLL(path)
        V = C.c_void_p
        S = C.c_size_t
        L.orp_brdf.argtypes = [C.c_int, V, _D, _D, _D, S, _D]
        L.orp_bsdf_pdf.argtypes = [C.c_int, V, _D, _D, _D, S, _D]
        L.orp_diffuse_bounce_weight.argtypes = [C.c_int, V, _D, _D, _D, S, _D]
        L.orp_lobe_probabilities.argtypes = [C.c_int, V, _D]
        L.orp_ggx_reflection_pdf.argtypes = [C.c_int, _D, C.c_double, C.c_double, _D, _D, S, _D]
        L.orp_ggx_d.argtypes = [C.c_int, C.c_int, _D, C.c_double, C.c_double, _D, S, _D]
        L.orp_ggx_g1.argtypes = [C.c_int, C.c_int, _D, C.c_double, C.c_double, _D, S, _D]
        L.orp_sample_vndf.argtypes = [C.c_int, _D, C.c_double, C.c_double, _D, _D, S, _D]
        L.orp_sample_ggx_half_vector.argtypes = [C.c_int, _D, C.c_double, _D, S, _D]
        L.orp_sample_bounce.argtypes = [C.c_int, V, _D, _D, C.c_uint32, S, _D, _D]
        L.orp_dielectric_fresnel.argtypes = [C.c_int, _D, C.c_double, C.c_double, S, _D]
        L.orp_sample_hg.argtypes = [C.c_int, C.c_double, _D, _D, S, _D]
        L.orp_sss_free_flight.argtypes = [C.c_int, _D, _D, S, _D]
        L.orp_sss_pdfs.argtypes = [C.c_int, _D, _D, S, _D, _D]
        L.orp_sss_weights.argtypes = [C.c_int, _D, _D, _D, S, _D, _D]
        L.orp_transmissive_surface.argtypes = [C.c_int, V, _D, C.c_int, C.c_double, _D, C.c_uint32, S, _D, _D, _I]
        L.orp_transmission_volume_walk.argtypes = [C.c_int, V, C.c_double, _D, C.c_uint32, S, _D, _I]
        self.L = L

    @staticmethod
    def _ok(rc):
        if rc != 0:
            raise ValueError("unsupported precision")

    @staticmethod
    def _mat(material) -> np.ndarray:
        m = np.ascontiguousarray(material)
        assert m.nbytes == 272, "a GPUMaterial record (272 B)"
        return m

    def brdf(self, material, frame, v, l, prec=64):
        m, fr = self._mat(material), _f64(frame)
        v, l = _f64(v, (-1, 3)), _f64(l, (-1, 3))
        out = np.empty((len(v), 3))
        self._ok(self.L.orp_brdf(prec, m.ctypes.data, _p(fr), _p(v), _p(l), len(v), _p(out)))
        return out

    def pdf(self, material, frame, v, l, prec=64):
        m, fr = self._mat(material), _f64(frame)
        v, l = _f64(v, (-1, 3)), _f64(l, (-1, 3))
        out = np.empty(len(v))
        self._ok(self.L.orp_bsdf_pdf(prec, m.ctypes.data, _p(fr), _p(v), _p(l), len(v), _p(out)))
        return out

    def diffuse_bounce_weight(self, material, frame, v, l, prec=64):
        m, fr = self._mat(material), _f64(frame)
        v, l = _f64(v, (-1, 3)), _f64(l, (-1, 3))
        out = np.empty((len(v), 3))
        self._ok(self.L.orp_diffuse_bounce_weight(prec, m.ctypes.data, _p(fr), _p(v), _p(l), len(v), _p(out)))
        return out

    def lobe_probabilities(self, material, prec=64):
        out = np.empty(3)
        self._ok(self.L.orp_lobe_probabilities(prec, self._mat(material).ctypes.data, _p(out)))
        return out  # diffuse, specular, clearcoat

    def ggx_pdf(self, frame, roughness, anisotropy, v, l, prec=64):
        fr, v, l = _f64(frame), _f64(v, (-1, 3)), _f64(l, (-1, 3))
        out = np.empty(len(v))
        self._ok(self.L.orp_ggx_reflection_pdf(prec, _p(fr), roughness, anisotropy, _p(v), _p(l), len(v), _p(out)))
        return out

    def ggx_d(self, frame, roughness, anisotropy, h, anisotropic=True, prec=64):
        fr, h = _f64(frame), _f64(h, (-1, 3))
        out = np.empty(len(h))
        self._ok(self.L.orp_ggx_d(prec, int(anisotropic), _p(fr), roughness, anisotropy, _p(h), len(h), _p(out)))
        return out

    def ggx_g1(self, frame, roughness, anisotropy, d, anisotropic=True, prec=64):
        fr, d = _f64(frame), _f64(d, (-1, 3))
        out = np.empty(len(d))
        self._ok(self.L.orp_ggx_g1(prec, int(anisotropic), _p(fr), roughness, anisotropy, _p(d), len(d), _p(out)))
        return out

    def sample_vndf(self, frame, roughness, anisotropy, v, u, prec=64):
        fr, v, u = _f64(frame), _f64(v), _f64(u, (-1, 2))
        out = np.empty((len(u), 3))
        self._ok(self.L.orp_sample_vndf(prec, _p(fr), roughness, anisotropy, _p(v), _p(u), len(u), _p(out)))
        return out

    def sample_ggx_half_vector(self, normal, roughness, u, prec=64):
        n, u = _f64(normal), _f64(u, (-1, 2))
        out = np.empty((len(u), 3))
        self._ok(self.L.orp_sample_ggx_half_vector(prec, _p(n), roughness, _p(u), len(u), _p(out)))
        return out

    def sample_bounce(self, material, frame, incoming, n, seed=1, prec=64):
        m, fr, inc = self._mat(material), _f64(frame), _f64(incoming)
        d, w = np.empty((n, 3)), np.empty((n, 3))
        self._ok(self.L.orp_sample_bounce(prec, m.ctypes.data, _p(fr), _p(inc), seed, n, _p(d), _p(w)))
        return d, w

    def dielectric_fresnel(self, cos_i, eta_i, eta_t, prec=64):
        c = _f64(cos_i, (-1,))
        out = np.empty(len(c))
        self._ok(self.L.orp_dielectric_fresnel(prec, _p(c), eta_i, eta_t, len(c), _p(out)))
        return out

    def sample_hg(self, g, direction, u, prec=64):
        d, u = _f64(direction), _f64(u, (-1, 2))
        out = np.empty((len(u), 3))
        self._ok(self.L.orp_sample_hg(prec, g, _p(d), _p(u), len(u), _p(out)))
        return out

    def free_flight(self, sigma_t, u, prec=64):
        s, u = _f64(sigma_t), _f64(u, (-1, 2))
        out = np.empty(len(u))
        self._ok(self.L.orp_sss_free_flight(prec, _p(s), _p(u), len(u), _p(out)))
        return out

    def sss_pdfs(self, sigma_t, d, prec=64):
        s, d = _f64(sigma_t), _f64(d, (-1,))
        coll, bound = np.empty(len(d)), np.empty(len(d))
        self._ok(self.L.orp_sss_pdfs(prec, _p(s), _p(d), len(d), _p(coll), _p(bound)))
        return coll, bound

    def sss_weights(self, sigma_t, albedo, d, prec=64):
        s, a, d = _f64(sigma_t), _f64(albedo), _f64(d, (-1,))
        coll, bound = np.empty((len(d), 3)), np.empty((len(d), 3))
        self._ok(self.L.orp_sss_weights(prec, _p(s), _p(a), _p(d), len(d), _p(coll), _p(bound)))
        return coll, bound

    def transmissive_surface(self, material, normal, front_facing, t, incoming, n, seed=1, prec=64):
        m, nn, inc = self._mat(material), _f64(normal), _f64(incoming)
        d, w, taken = np.empty((n, 3)), np.empty((n, 3)), np.empty(n, np.int32)
        self._ok(self.L.orp_transmissive_surface(prec, m.ctypes.data, _p(nn), int(front_facing), t, _p(inc), seed, n,
                                                 _p(d), _p(w), taken.ctypes.data_as(_I)))
        return d, w, taken.astype(bool)

    def transmission_volume_walk(self, material, t, incoming, n, seed=1, prec=64):
        m, inc = self._mat(material), _f64(incoming)
        w, sc = np.empty((n, 3)), np.empty(n, np.int32)
        self._ok(self.L.orp_transmission_volume_walk(prec, m.ctypes.data, t, _p(inc), seed, n, _p(w),
                                                     sc.ctypes.data_as(_I)))
        return w, sc  # 0 escaped (or not a volume), 1 scattered, 2 scattered and ended (throughput <= 1e-4)
