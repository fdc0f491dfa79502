"""SceneScript: the reference's line-based ``.denrim`` scene language
(Scripting/SceneScript.swift), read into a ``RenderScene`` for this backend.

    scene = parse_file("material-testball.denrim")
    settings, samples = scene.render_defaults.render_settings(), scene.render_defaults.samples

Commands (SceneScript.swift:125-220): ``include``, ``render``/``defaults``,
``camera``, ``environment`` (``sky`` or an HDR/image), ``texture`` (``solid``,
``checker``, ``image``, ``normalFrom``), ``mesh`` (OBJ/PLY, optional
``flipV``), ``material`` (RGB, RGB + emission, or ``preset``, then keyword
parameters), ``quad``, ``box``, ``instance`` and ``sdf``/``volume``.
Lines are cut at ``#``; ``(``, ``)`` and ``,`` separate tokens; keywords are
case-insensitive; numbers read as Swift's ``Float(String)`` (correctly
rounded). Errors are ``SceneScriptError`` with the reference's cases and
messages (SceneScript.swift:5-46).

``sdf`` builds the field with this backend's volume builders (dense grid or
sparse bricks, the reference's ``.cpuReference`` bake at the given cubic
resolution). Per-primitive material fields (``color``, ``opacity``,
``emission``, ``roughness``, ``metallic``, ``transmission``) and custom
attribute channels (``attr``/``attributes``) feed the reference's material
programs, which this backend does not have: a script that sets them raises
``SceneScriptError`` (kind ``unsupported``) instead of rendering something
else.
"""
from __future__ import annotations

import os
from dataclasses import dataclass
from typing import Callable, Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import materials as _presets
from .mesh_import import MeshLoadingError, load_mesh, parse_float32, parse_int
from .scene import (Camera, Environment, Material, Mesh, RenderScene, Texture2D, TextureSamplingMode, Transform,
                    compile_scene)
from .settings import DenoiseSettings, RenderAccelerationMode, RenderOutput, RenderQuality, RenderSettings
from .texture_io import TextureLoadingError, derived_normal_map, load_texture

f32 = np.float32


class SceneScriptError(Exception):
    """SceneScriptError (SceneScript.swift:5-46)."""

    _MESSAGES = {
        "unknownCommand": "Unknown scene script command '{v}' on line {line}.",
        "invalidArgumentCount": "Invalid argument count for scene script command '{v}' on line {line}.",
        "invalidNumber": "Invalid numeric value '{v}' on line {line}.",
        "unknownMaterial": "Unknown material '{v}' on line {line}.",
        "unknownMaterialPreset": "Unknown built-in material preset '{v}' on line {line}.",
        "unknownTexture": "Unknown texture '{v}' on line {line}.",
        "unknownMesh": "Unknown mesh '{v}' on line {line}.",
        "environmentLoadFailed": "Could not load environment image '{v}' on line {line}.",
        "textureLoadFailed": "Could not load texture image '{v}' on line {line}.",
        "meshLoadFailed": "Could not load mesh '{v}' on line {line}.",
        "includeResolverMissing": "Scene script include '{v}' on line {line} requires an include resolver.",
        "includeCycle": "Scene script include cycle for '{v}' on line {line}.",
        "unsupported": "Scene script feature '{v}' on line {line} needs material programs, which this backend "
                       "does not have.",
    }

    def __init__(self, kind: str, value: str, line: int):
        self.kind, self.value, self.line = kind, value, line
        super().__init__(self._MESSAGES[kind].format(v=value, line=line))

    def __eq__(self, other):
        return isinstance(other, SceneScriptError) and (self.kind, self.value, self.line) == (
            other.kind, other.value, other.line)

    def __hash__(self):
        return hash((self.kind, self.value, self.line))


@dataclass
class RenderDefaults:
    """RenderDefaults (Scene/RenderScene.swift:5-67): script-authored render
    options; None = not set."""
    output_path: Optional[str] = None
    output: Optional[RenderOutput] = None
    samples: Optional[int] = None
    width: Optional[int] = None
    height: Optional[int] = None
    quality: Optional[RenderQuality] = None
    max_bounces: Optional[int] = None
    acceleration_mode: Optional[RenderAccelerationMode] = None
    sample_radiance_clamp: Optional[float] = None
    transparent_background: Optional[bool] = None
    denoise: Optional[DenoiseSettings] = None
    sdf_resolution: Optional[int] = None

    def render_settings(self, base: Optional[RenderSettings] = None) -> RenderSettings:
        """`base` (RenderSettings() if None) with the fields the script set."""
        s = base or RenderSettings()
        kw = {}
        for name, attr in (("width", "width"), ("height", "height"), ("quality", "quality"),
                           ("max_bounces", "max_bounces"), ("sample_radiance_clamp", "sample_radiance_clamp"),
                           ("transparent_background", "transparent_background"), ("denoise", "denoise")):
            v = getattr(self, name)
            if v is not None:
                kw[attr] = v
        import dataclasses
        return dataclasses.replace(s, **kw)


IncludeResolver = Callable[[str], str]


# ------------------------------------------------------------ token helpers --
def _strip_comment(line: str) -> str:
    k = line.find("#")
    return line if k < 0 else line[:k]


def _tokenize(line: str) -> List[str]:
    return line.replace("(", " ( ").replace(")", " ) ").replace(",", " ").split()


class _Ctx:
    def __init__(self, base_dir, include_resolver, sdf_override):
        self.base_dir = base_dir
        self.include_resolver = include_resolver
        self.sdf_override = sdf_override
        self.materials: Dict[str, int] = {}
        self.textures: Dict[str, Texture2D] = {}
        self.meshes: Dict[str, Mesh] = {}
        self.include_stack: List[str] = []


def _floats(tokens: Sequence[str], line: int) -> List[float]:
    out = []
    for t in tokens:
        v = parse_float32(t)
        if v is None:
            raise SceneScriptError("invalidNumber", t, line)
        out.append(v)
    return out


def _float1(tokens, i, line) -> float:
    return _floats([tokens[i]], line)[0]


def _int(token: str, line: int) -> int:
    v = parse_int(token)
    if v is None:
        raise SceneScriptError("invalidNumber", token, line)
    return v


def _bool(token: str, line: int) -> bool:
    t = token.lower()
    if t in ("1", "true", "yes", "on"):
        return True
    if t in ("0", "false", "no", "off"):
        return False
    raise SceneScriptError("invalidNumber", token, line)


def _can_parse_floats(tokens) -> bool:
    return all(parse_float32(t) is not None for t in tokens)


def _named_vector(tokens, i, line, n) -> Tuple[str, List[float], int]:
    """parseNamedVector2/3 (SceneScript.swift:1603-1708): ``name(a, b, c)``
    or ``name a b c``."""
    if i >= len(tokens):
        raise SceneScriptError("invalidArgumentCount", "vector", line)
    name = tokens[i]
    if i + 1 < len(tokens) and tokens[i + 1] == "(":
        if not (i + n + 2 < len(tokens)) or tokens[i + n + 2] != ")":
            raise SceneScriptError("invalidArgumentCount", name, line)
        return name, _floats(tokens[i + 2:i + 2 + n], line), i + n + 3
    if i + n >= len(tokens):
        raise SceneScriptError("invalidArgumentCount", name, line)
    return name, _floats(tokens[i + 1:i + 1 + n], line), i + n + 1


def _named_float(tokens, i, line) -> Tuple[str, float, int]:
    """parseNamedFloat (SceneScript.swift:1710-1741)."""
    if i >= len(tokens):
        raise SceneScriptError("invalidArgumentCount", "float", line)
    name = tokens[i]
    if i + 1 < len(tokens) and tokens[i + 1] == "(":
        if not (i + 3 < len(tokens)) or tokens[i + 3] != ")":
            raise SceneScriptError("invalidArgumentCount", name, line)
        return name, _float1(tokens, i + 2, line), i + 4
    if i + 1 >= len(tokens):
        raise SceneScriptError("invalidArgumentCount", name, line)
    return name, _float1(tokens, i + 1, line), i + 2


def _named_matrix4(tokens, i, line) -> Tuple[str, np.ndarray, int]:
    """parseNamedMatrix4 (SceneScript.swift:1638-1673): 16 values, row-major."""
    if i >= len(tokens):
        raise SceneScriptError("invalidArgumentCount", "matrix", line)
    name = tokens[i]
    if i + 1 < len(tokens) and tokens[i + 1] == "(":
        if not (i + 18 < len(tokens)) or tokens[i + 18] != ")":
            raise SceneScriptError("invalidArgumentCount", name, line)
        vals, nxt = _floats(tokens[i + 2:i + 18], line), i + 19
    else:
        if not (i + 16 < len(tokens)):
            raise SceneScriptError("invalidArgumentCount", name, line)
        vals, nxt = _floats(tokens[i + 1:i + 17], line), i + 17
    return name, np.array(vals, f32).reshape(4, 4), nxt


def _optional_named_string(tokens, i, keyword, bare_ok, line) -> Tuple[str, int]:
    """parseOptionalNamedString (SceneScript.swift:1570-1601)."""
    if i >= len(tokens):
        raise SceneScriptError("invalidArgumentCount", keyword, line)
    if tokens[i].lower() == keyword.lower():
        if i + 1 < len(tokens) and tokens[i + 1] == "(":
            if not (i + 3 < len(tokens)) or tokens[i + 3] != ")":
                raise SceneScriptError("invalidArgumentCount", keyword, line)
            return tokens[i + 2], i + 4
        if i + 1 >= len(tokens):
            raise SceneScriptError("invalidArgumentCount", keyword, line)
        return tokens[i + 1], i + 2
    if not bare_ok:
        raise SceneScriptError("invalidArgumentCount", keyword, line)
    return tokens[i], i + 1


def _asset_path(path: str, base_dir: Optional[str]) -> str:
    if path.startswith("/"):
        return path
    return os.path.join(base_dir if base_dir is not None else os.getcwd(), path)


def _material_id(name, line, ctx) -> int:
    if name not in ctx.materials:
        raise SceneScriptError("unknownMaterial", name, line)
    return ctx.materials[name]


def _texture(name, line, ctx) -> Texture2D:
    if name not in ctx.textures:
        raise SceneScriptError("unknownTexture", name, line)
    return ctx.textures[name]


# ---------------------------------------------------------------- commands --
_RENDER_KEYWORDS = {"output", "path", "outputtype", "output-type", "aov", "samples", "sample", "spp", "size",
                    "resolution", "width", "height", "quality", "maxbounces", "max-bounces", "bounces", "backend",
                    "acceleration", "sampleradianceclamp", "sample-radiance-clamp", "clamp",
                    "transparentbackground", "transparent-background", "denoise", "sdfresolution",
                    "sdf-resolution", "fieldresolution", "field-resolution", "volumeresolution",
                    "volume-resolution"}


def _render_output(name, line) -> RenderOutput:
    m = {"beauty": RenderOutput.BEAUTY, "depth": RenderOutput.DEPTH, "normal": RenderOutput.NORMAL,
         "normals": RenderOutput.NORMAL, "albedo": RenderOutput.ALBEDO, "material-id": RenderOutput.MATERIAL_ID,
         "materialid": RenderOutput.MATERIAL_ID, "object-id": RenderOutput.OBJECT_ID,
         "objectid": RenderOutput.OBJECT_ID, "motion": RenderOutput.MOTION_VECTOR,
         "motion-vector": RenderOutput.MOTION_VECTOR, "motionvector": RenderOutput.MOTION_VECTOR}
    if name.lower() not in m:
        raise SceneScriptError("invalidArgumentCount", "render", line)
    return m[name.lower()]


def _render_quality(name, line) -> RenderQuality:
    m = {"preview": RenderQuality.PREVIEW, "fast": RenderQuality.PREVIEW, "interactive": RenderQuality.INTERACTIVE,
         "viewport": RenderQuality.INTERACTIVE, "final": RenderQuality.FINAL, "export": RenderQuality.FINAL}
    if name.lower() not in m:
        raise SceneScriptError("invalidArgumentCount", "render", line)
    return m[name.lower()]


def _render_backend(name, line) -> RenderAccelerationMode:
    n = name.lower()
    if n in ("automatic", "auto"):
        return RenderAccelerationMode.AUTOMATIC
    if n in ("flat", "flat-bvh", "flatbvh"):
        return RenderAccelerationMode.FLAT_BVH
    if n in ("metal", "metal-ray-tracing", "metalrt", "hardware", "hardware-ray-tracing"):
        return RenderAccelerationMode.METAL_RAY_TRACING
    if n == "hip":
        return RenderAccelerationMode.HIP
    raise SceneScriptError("invalidArgumentCount", "render", line)


def _denoise(name, line) -> DenoiseSettings:
    n = name.lower()
    if n in ("none", "off", "false"):
        return DenoiseSettings.none()
    if n in ("simple", "simple-spatial", "spatial", "experimental-simple"):
        return DenoiseSettings.simple_spatial()
    if n in ("apple", "mps", "svgf", "apple-svgf"):
        return DenoiseSettings.apple_svgf()
    if n == "learned":
        return DenoiseSettings.learned()
    raise SceneScriptError("invalidArgumentCount", "render", line)


def _render_size(name, line) -> Tuple[int, int]:
    n = name.lower()
    if n in ("hd", "720p"):
        return 1280, 720
    if n in ("fhd", "fullhd", "1080p"):
        return 1920, 1080
    if n in ("uhd", "4k", "2160p"):
        return 3840, 2160
    s = max(1, _int(name, line))
    return s, s


def _parse_render_defaults(tokens, line, base_dir, cur: RenderDefaults) -> RenderDefaults:
    """parseRenderDefaults (SceneScript.swift:245-347)."""
    import dataclasses
    d = dataclasses.replace(cur)
    i = 1
    if i < len(tokens) and tokens[i].lower() == "defaults":
        i += 1

    def need():
        if i + 1 >= len(tokens):
            raise SceneScriptError("invalidArgumentCount", "render", line)

    while i < len(tokens):
        k = tokens[i].lower()
        if k in ("output", "path"):
            need()
            p = tokens[i + 1]
            d.output_path = os.path.normpath(p if p.startswith("/") or base_dir is None else os.path.join(base_dir, p))
            i += 2
        elif k in ("outputtype", "output-type", "aov"):
            need()
            d.output = _render_output(tokens[i + 1], line)
            i += 2
        elif k in ("samples", "sample", "spp"):
            need()
            d.samples = max(0, _int(tokens[i + 1], line))
            i += 2
        elif k in ("size", "resolution"):
            need()
            d.width, d.height = _render_size(tokens[i + 1], line)
            i += 2
        elif k == "width":
            need()
            d.width = max(1, _int(tokens[i + 1], line))
            i += 2
        elif k == "height":
            need()
            d.height = max(1, _int(tokens[i + 1], line))
            i += 2
        elif k == "quality":
            need()
            d.quality = _render_quality(tokens[i + 1], line)
            i += 2
        elif k in ("maxbounces", "max-bounces", "bounces"):
            need()
            d.max_bounces = max(1, _int(tokens[i + 1], line))
            i += 2
        elif k in ("backend", "acceleration"):
            need()
            d.acceleration_mode = _render_backend(tokens[i + 1], line)
            i += 2
        elif k in ("sampleradianceclamp", "sample-radiance-clamp", "clamp"):
            need()
            d.sample_radiance_clamp = _float1(tokens, i + 1, line)
            i += 2
        elif k in ("transparentbackground", "transparent-background"):
            if i + 1 < len(tokens) and tokens[i + 1].lower() not in _RENDER_KEYWORDS:
                d.transparent_background = _bool(tokens[i + 1], line)
                i += 2
            else:
                d.transparent_background = True
                i += 1
        elif k == "denoise":
            need()
            d.denoise = _denoise(tokens[i + 1], line)
            i += 2
        elif k in ("sdfresolution", "sdf-resolution", "fieldresolution", "field-resolution", "volumeresolution",
                   "volume-resolution"):
            need()
            d.sdf_resolution = max(2, _int(tokens[i + 1], line))
            i += 2
        else:
            raise SceneScriptError("invalidArgumentCount", "render", line)
    return d


def _parse_camera(tokens, line) -> Camera:
    """parseCamera (SceneScript.swift:349-394)."""
    if len(tokens) == 8:
        v = _floats(tokens[1:], line)
        return Camera(origin=v[0:3], target=v[3:6], vertical_field_of_view_degrees=v[6])
    origin = target = fov = None
    i = 1
    while i < len(tokens):
        k = tokens[i].lower()
        if k in ("origin", "position", "eye"):
            _, origin, i = _named_vector(tokens, i, line, 3)
        elif k in ("target", "lookat", "look"):
            _, target, i = _named_vector(tokens, i, line, 3)
        elif k in ("fov", "fieldofview"):
            _, fov, i = _named_float(tokens, i, line)
        else:
            raise SceneScriptError("invalidArgumentCount", "camera", line)
    if origin is None or target is None or fov is None:
        raise SceneScriptError("invalidArgumentCount", "camera", line)
    return Camera(origin=origin, target=target, vertical_field_of_view_degrees=fov)


def _sampling_mode(rest, line) -> int:
    if not rest:
        return TextureSamplingMode.NEAREST
    if len(rest) != 1:
        raise SceneScriptError("invalidArgumentCount", "texture", line)
    m = rest[0].lower()
    if m == "nearest":
        return TextureSamplingMode.NEAREST
    if m == "linear":
        return TextureSamplingMode.LINEAR
    raise SceneScriptError("invalidArgumentCount", "texture", line)


def _rgba(tokens, line) -> List[float]:
    v = _floats(tokens, line)
    if len(v) != 4:
        raise SceneScriptError("invalidArgumentCount", "texture", line)
    return v


def _parse_texture(tokens, line, ctx) -> Tuple[str, Texture2D]:
    """parseTexture (SceneScript.swift:396-466)."""
    if len(tokens) < 4:
        raise SceneScriptError("invalidArgumentCount", "texture", line)
    name, kind = tokens[1], tokens[2].lower()
    if kind == "solid":
        if len(tokens) not in (7, 8):
            raise SceneScriptError("invalidArgumentCount", "texture", line)
        c = _rgba(tokens[3:7], line)
        return name, Texture2D(1, 1, np.array([c], f32), _sampling_mode(tokens[7:], line))
    if kind == "checker":
        if len(tokens) not in (11, 12):
            raise SceneScriptError("invalidArgumentCount", "texture", line)
        a, b = _rgba(tokens[3:7], line), _rgba(tokens[7:11], line)
        # Texture2D.checker (Texture.swift:152-163): 2x2, a b / b a.
        return name, Texture2D(2, 2, np.array([a, b, b, a], f32), _sampling_mode(tokens[11:], line))
    if kind == "image":
        enc, mode = "srgb", TextureSamplingMode.LINEAR
        rest = tokens[4:]
        j = 0
        while j < len(rest):
            k = rest[j].lower()
            if k == "color":
                if j + 1 >= len(rest) or rest[j + 1].lower() not in ("srgb", "linear"):
                    raise SceneScriptError("invalidArgumentCount", "texture", line)
                enc = rest[j + 1].lower()
                j += 2
            elif k == "sampler":
                if j + 1 >= len(rest):
                    raise SceneScriptError("invalidArgumentCount", "texture", line)
                mode = _sampling_mode([rest[j + 1]], line)
                j += 2
            else:
                raise SceneScriptError("invalidArgumentCount", "texture", line)
        try:
            return name, load_texture(_asset_path(tokens[3], ctx.base_dir), enc, mode)
        except (TextureLoadingError, OSError):
            raise SceneScriptError("textureLoadFailed", tokens[3], line)
    if kind in ("normalfrom", "normalfromtexture"):
        if len(tokens) not in (4, 6):
            raise SceneScriptError("invalidArgumentCount", "texture", line)
        strength = 1.0
        if len(tokens) == 6:
            if tokens[4].lower() != "strength":
                raise SceneScriptError("invalidArgumentCount", "texture", line)
            strength = _float1(tokens, 5, line)
        return name, derived_normal_map(_texture(tokens[3], line, ctx), strength)
    raise SceneScriptError("invalidArgumentCount", "texture", line)


def _parse_environment(tokens, line, ctx) -> Environment:
    """parseEnvironment (SceneScript.swift:468-541)."""
    if len(tokens) < 2:
        raise SceneScriptError("invalidArgumentCount", "environment", line)
    kind = tokens[1].lower()
    if kind in ("sky", "default"):
        if len(tokens) != 2:
            raise SceneScriptError("invalidArgumentCount", "environment", line)
        return Environment()
    if kind not in ("image", "hdri", "hdr") or len(tokens) < 3:
        raise SceneScriptError("invalidArgumentCount", "environment", line)
    intensity, rot, maxr = 1.0, 0.0, 16.0
    i = 3
    while i < len(tokens):
        k = tokens[i].lower()
        if i + 1 >= len(tokens):
            raise SceneScriptError("invalidArgumentCount", "environment", line)
        if k in ("intensity", "strength", "exposure"):
            intensity = _float1(tokens, i + 1, line)
        elif k in ("rotationy", "rotatey", "yaw"):
            rot = _float1(tokens, i + 1, line)
        elif k in ("maxradiance", "radianceclamp", "clamp"):
            maxr = _float1(tokens, i + 1, line)
        else:
            raise SceneScriptError("invalidArgumentCount", "environment", line)
        i += 2
    try:
        tex = load_texture(_asset_path(tokens[2], ctx.base_dir), "linear", TextureSamplingMode.LINEAR)
    except (TextureLoadingError, OSError):
        raise SceneScriptError("environmentLoadFailed", tokens[2], line)
    return Environment(texture=tex, intensity=intensity, rotation_y=rot, max_radiance=maxr)


def _parse_mesh(tokens, line, ctx) -> Tuple[str, Mesh]:
    """parseMeshAsset (SceneScript.swift:543-583); flipV maps v -> 1 - v."""
    if len(tokens) not in (3, 4):
        raise SceneScriptError("invalidArgumentCount", "mesh", line)
    flip = False
    if len(tokens) == 4:
        if tokens[3].lower() != "flipv":
            raise SceneScriptError("invalidArgumentCount", "mesh", line)
        flip = True
    try:
        m = load_mesh(_asset_path(tokens[2], ctx.base_dir))
    except (MeshLoadingError, OSError):
        raise SceneScriptError("meshLoadFailed", tokens[2], line)
    if flip and len(m.texcoords):
        m.texcoords = m.texcoords.copy()
        m.texcoords[:, 1] = f32(1) - m.texcoords[:, 1]
    return tokens[1], m


# Material keywords: (names, attribute, component count) (SceneScript.swift:631-882).
_MATERIAL_KEYS = [
    (("roughness",), "roughness", 1), (("metallic",), "metallic", 1), (("specular",), "specular", 1),
    (("specularcolor",), "specular_color", 3), (("ior", "indexofrefraction"), "index_of_refraction", 1),
    (("anisotropy", "specularanisotropy"), "specular_anisotropy", 1), (("clearcoat",), "clearcoat", 1),
    (("clearcoatcolor", "clearcoattint"), "clearcoat_color", 3),
    (("clearcoatattenuationcolor", "clearcoatabsorptioncolor"), "clearcoat_attenuation_color", 3),
    (("clearcoatthickness", "clearcoatdepth"), "clearcoat_thickness", 1),
    (("clearcoatroughness",), "clearcoat_roughness", 1),
    (("clearcoatior", "clearcoatindexofrefraction"), "clearcoat_index_of_refraction", 1),
    (("thinfilm", "iridescence"), "thin_film", 1),
    (("thinfilmthickness", "thinfilmthicknessnm", "iridescencethickness"), "thin_film_thickness_nanometers", 1),
    (("thinfilmior", "thinfilmindexofrefraction", "iridescenceior"), "thin_film_index_of_refraction", 1),
    (("sheen", "fuzz"), "sheen", 1), (("sheencolor", "fuzzcolor"), "sheen_color", 3),
    (("sheenroughness", "fuzzroughness"), "sheen_roughness", 1),
    (("subsurface", "subsurfaceweight", "sss"), "subsurface", 1),
    (("subsurfacecolor", "subsurfacetint", "ssscolor", "scatteringcolor"), "subsurface_color", 3),
    (("subsurfaceradius", "subsurfaceradiuscolor", "sssradius", "scatteringradius"), "subsurface_radius", 3),
    (("subsurfacescale", "sssscale", "scatteringscale"), "subsurface_scale", 1),
    (("subsurfaceanisotropy", "sssanisotropy", "scatteringanisotropy"), "subsurface_anisotropy", 1),
    (("opacity",), "opacity", 1),
    (("transmission", "spectrans", "spectraltransmission"), "transmission", 1),
    (("transmissioncolor", "transmissiontint", "spectranscolor", "spectraltransmissioncolor"),
     "transmission_color", 3),
    (("transmissionroughness", "spectransroughness", "spectraltransmissionroughness"), "transmission_roughness", 1),
    (("transmissionior", "transmissionindexofrefraction"), "transmission_index_of_refraction", 1),
    (("transmissionabsorptioncolor", "absorptioncolor", "attenuationcolor"), "transmission_absorption_color", 3),
    (("transmissionabsorptiondistance", "absorptiondistance", "attenuationdistance"),
     "transmission_absorption_distance", 1),
    (("volumescattering", "volumescatter", "mediumscattering", "mediumscatter"), "volume_scattering", 1),
    (("volumescatteringcolor", "volumescattercolor", "mediumscatteringcolor", "mediumscattercolor"),
     "volume_scattering_color", 3),
    (("volumescatteringdistance", "volumescatterdistance", "mediumscatteringdistance", "mediumscatterdistance"),
     "volume_scattering_distance", 1),
    (("volumeanisotropy", "mediumanisotropy"), "volume_anisotropy", 1),
]
_MATERIAL_LOOKUP = {n: (attr, k) for names, attr, k in _MATERIAL_KEYS for n in names}


def _parse_material(tokens, line, ctx) -> Tuple[str, Material]:
    """parseMaterial (SceneScript.swift:585-890)."""
    import dataclasses
    if len(tokens) < 3:
        raise SceneScriptError("invalidArgumentCount", "material", line)
    name, src = tokens[1], tokens[2].lower()
    if src in ("preset", "builtin", "built-in"):
        if len(tokens) < 4:
            raise SceneScriptError("invalidArgumentCount", "material", line)
        preset = _presets.material(tokens[3])
        if preset is None:
            raise SceneScriptError("unknownMaterialPreset", tokens[3], line)
        m, i = dataclasses.replace(preset), 4
    else:
        if len(tokens) < 5:
            raise SceneScriptError("invalidArgumentCount", "material", line)
        base = _floats(tokens[2:5], line)
        if len(tokens) == 9 and _can_parse_floats(tokens[5:9]):
            e = _floats(tokens[5:9], line)
            return name, Material(base_color=tuple(base), emission=tuple(e[:3]), emission_strength=e[3])
        m, i = Material(base_color=tuple(base)), 5
    while i < len(tokens):
        k = tokens[i].lower()
        if k == "emission":
            if i + 4 >= len(tokens):
                raise SceneScriptError("invalidArgumentCount", "material", line)
            v = _floats(tokens[i + 1:i + 5], line)
            m.emission, m.emission_strength = tuple(v[:3]), v[3]
            i += 5
        elif k in ("thinwalled", "thinwall", "thin"):
            if i + 1 >= len(tokens):
                raise SceneScriptError("invalidArgumentCount", "material", line)
            m.thin_walled = _float1(tokens, i + 1, line) > 0
            i += 2
        elif k in ("basecolortexture", "normalmap"):
            if i + 1 >= len(tokens):
                raise SceneScriptError("invalidArgumentCount", "material", line)
            t = _texture(tokens[i + 1], line, ctx)
            if k == "basecolortexture":
                m.base_color_texture = t
            else:
                m.normal_map = t
            i += 2
        elif k in _MATERIAL_LOOKUP:
            attr, n = _MATERIAL_LOOKUP[k]
            if i + n >= len(tokens):
                raise SceneScriptError("invalidArgumentCount", "material", line)
            v = _floats(tokens[i + 1:i + 1 + n], line)
            setattr(m, attr, v[0] if n == 1 else tuple(v))
            i += 1 + n
        else:
            raise SceneScriptError("invalidArgumentCount", "material", line)
    return name, m


def _parse_quad(tokens, line, ctx) -> Tuple[Mesh, int]:
    """parseQuad (SceneScript.swift:892-959)."""
    if len(tokens) == 14:
        mat = _material_id(tokens[1], line, ctx)
        v = _floats(tokens[2:14], line)
        return Mesh.quad(v[0:3], v[3:6], v[6:9], v[9:12]), mat
    if len(tokens) < 2:
        raise SceneScriptError("invalidArgumentCount", "quad", line)
    mname, i = _optional_named_string(tokens, 1, "material", True, line)
    mat = _material_id(mname, line, ctx)
    corners, uvs = {}, {}
    while i < len(tokens):
        k = tokens[i].lower()
        if k in ("uva", "uvb", "uvc", "uvd"):
            _, uvs[k], i = _named_vector(tokens, i, line, 2)
        else:
            nm, v, i = _named_vector(tokens, i, line, 3)
            corners[nm.lower()] = v
    if not all(c in corners for c in "abcd"):
        raise SceneScriptError("invalidArgumentCount", "quad", line)
    if len(uvs) == 4:
        tc = [uvs["uva"], uvs["uvb"], uvs["uvc"], uvs["uvd"]]
    elif not uvs:
        tc = [(0, 0), (1, 0), (1, 1), (0, 1)]
    else:
        raise SceneScriptError("invalidArgumentCount", "quad", line)
    return Mesh.quad(corners["a"], corners["b"], corners["c"], corners["d"], texcoords=tc), mat


def _parse_box(tokens, line, ctx) -> Tuple[Mesh, int, Transform]:
    """parseBox (SceneScript.swift:961-1025)."""
    if len(tokens) in (8, 9):
        mat = _material_id(tokens[1], line, ctx)
        v = _floats(tokens[2:], line)
        t = Transform.translation(v[0:3])
        if len(v) == 7:
            t = t * Transform.rotation_y(v[6])
        return Mesh.box_sized(v[3:6]), mat, t
    mname, i = _optional_named_string(tokens, 1, "material", True, line)
    mat = _material_id(mname, line, ctx)
    pos, size, rot = [0.0, 0.0, 0.0], None, None
    while i < len(tokens):
        k = tokens[i].lower()
        if k in ("position", "translation", "translate"):
            _, pos, i = _named_vector(tokens, i, line, 3)
        elif k in ("size", "scale"):
            _, size, i = _named_vector(tokens, i, line, 3)
        elif k in ("rotationy", "rotatey"):
            _, rot, i = _named_float(tokens, i, line)
        else:
            raise SceneScriptError("invalidArgumentCount", "box", line)
    if size is None:
        raise SceneScriptError("invalidArgumentCount", "box", line)
    t = Transform.translation(pos)
    if rot is not None:
        t = t * Transform.rotation_y(rot)
    return Mesh.box_sized(size), mat, t


def _parse_instance(tokens, line, ctx) -> Tuple[Mesh, int, Transform]:
    """parseInstance (SceneScript.swift:1027-1112)."""
    if len(tokens) in (9, 10):
        if tokens[1] not in ctx.meshes:
            raise SceneScriptError("unknownMesh", tokens[1], line)
        mesh = ctx.meshes[tokens[1]]
        mat = _material_id(tokens[2], line, ctx)
        v = _floats(tokens[3:], line)
        t = Transform.translation(v[0:3]) * Transform.scale(v[3:6])
        if len(v) == 7:
            t = Transform.translation(v[0:3]) * Transform.rotation_y(v[6]) * Transform.scale(v[3:6])
        return mesh, mat, t
    mesh_name, i = _optional_named_string(tokens, 1, "mesh", True, line)
    mat_name, i = _optional_named_string(tokens, i, "material", True, line)
    if mesh_name not in ctx.meshes:
        raise SceneScriptError("unknownMesh", mesh_name, line)
    mesh = ctx.meshes[mesh_name]
    mat = _material_id(mat_name, line, ctx)
    pos, scale = [0.0, 0.0, 0.0], [1.0, 1.0, 1.0]
    rx = ry = rz = None
    while i < len(tokens):
        k = tokens[i].lower()
        if k in ("position", "translation", "translate"):
            _, pos, i = _named_vector(tokens, i, line, 3)
        elif k == "scale":
            _, scale, i = _named_vector(tokens, i, line, 3)
        elif k in ("rotationx", "rotatex"):
            _, rx, i = _named_float(tokens, i, line)
        elif k in ("rotationy", "rotatey"):
            _, ry, i = _named_float(tokens, i, line)
        elif k in ("rotationz", "rotatez"):
            _, rz, i = _named_float(tokens, i, line)
        else:
            raise SceneScriptError("invalidArgumentCount", "instance", line)
    t = Transform.translation(pos)
    if rx is not None:
        t = t * Transform.rotation_x(rx)
    if ry is not None:
        t = t * Transform.rotation_y(ry)
    if rz is not None:
        t = t * Transform.rotation_z(rz)
    return mesh, mat, t * Transform.scale(scale)


# -------------------------------------------------------------------- sdf --
_SDF_SHAPES = ("sphere", "box", "cylinder")
_SDF_OPS = ("union", "add", "subtract", "difference", "cut")
_SDF_GLOBALS = ("dense", "sparse", "bricks", "brick", "backend", "storage", "material", "resolution", "res",
                "bricksize", "brick-size", "samplescale", "sample-scale", "sdfsamplescale", "sdf-sample-scale",
                "narrowband", "band", "boundsmin", "min", "boundsmax", "max", "attributes", "attrs", "position",
                "translation", "translate", "scale", "rotationx", "rotatex", "rotationy", "rotatey", "rotationz",
                "rotatez")
_SDF_PRIM_KEYS = ("material", "radius", "halfheight", "half-height", "height", "operation", "op", "subtract",
                  "difference", "cut", "union", "add", "halfextents", "halfextent", "extents", "size",
                  "cornerradius", "rounding", "round", "position", "translation", "translate", "scale",
                  "rotationx", "rotatex", "rotationy", "rotatey", "rotationz", "rotatez", "worldtolocal",
                  "worldtoprimitive", "world-to-local", "matrix", "smooth", "smoothunion", "smoothunionradius",
                  "attr", "attribute", "basecolor", "color", "opacity", "emission", "roughness", "metallic",
                  "transmission")


def _sdf_op(name, line) -> str:
    n = name.lower()
    if n in ("union", "add"):
        return "union"
    if n in ("subtract", "difference", "cut"):
        return "subtract"
    raise SceneScriptError("invalidArgumentCount", "sdf", line)


def _parse_sdf_primitive(tokens, i, line, fallback, ctx):
    """parseSDFPrimitive (SceneScript.swift:1306-1535)."""
    from .sdf import SDFPrimitive
    shape = tokens[i].lower()
    op = "union"
    i += 1
    if shape in _SDF_OPS:
        op = _sdf_op(shape, line)
        if i >= len(tokens) or tokens[i].lower() not in _SDF_SHAPES:
            raise SceneScriptError("invalidArgumentCount", "sdf", line)
        shape = tokens[i].lower()
        i += 1
    mname = fallback
    if i < len(tokens) and tokens[i].lower() not in _SDF_PRIM_KEYS and tokens[i].lower() not in _SDF_SHAPES \
            and tokens[i].lower() not in _SDF_OPS:
        mname = tokens[i]
        i += 1
    radius = half_height = half_ext = None
    corner, smooth = 0.0, 0.0
    pos, scale = [0.0, 0.0, 0.0], [1.0, 1.0, 1.0]
    rx = ry = rz = None
    w2l = None
    while i < len(tokens) and tokens[i].lower() not in _SDF_SHAPES and tokens[i].lower() not in _SDF_OPS:
        k = tokens[i].lower()

        def one():
            if i + 1 >= len(tokens):
                raise SceneScriptError("invalidArgumentCount", "sdf", line)
            return _float1(tokens, i + 1, line)

        if k == "material":
            mname, i = _optional_named_string(tokens, i, "material", False, line)
        elif k == "radius":
            radius, i = one(), i + 2
        elif k in ("halfheight", "half-height"):
            half_height, i = one(), i + 2
        elif k == "height":
            half_height, i = one() * 0.5, i + 2
        elif k in ("operation", "op"):
            if i + 1 >= len(tokens):
                raise SceneScriptError("invalidArgumentCount", "sdf", line)
            op, i = _sdf_op(tokens[i + 1], line), i + 2
        elif k in ("subtract", "difference", "cut"):
            op, i = "subtract", i + 1
        elif k in ("union", "add"):
            op, i = "union", i + 1
        elif k in ("halfextents", "halfextent", "extents"):
            _, half_ext, i = _named_vector(tokens, i, line, 3)
        elif k == "size":
            _, v, i = _named_vector(tokens, i, line, 3)
            half_ext = [float(f32(x) * f32(0.5)) for x in v]
        elif k in ("cornerradius", "rounding", "round"):
            corner, i = one(), i + 2
        elif k in ("position", "translation", "translate"):
            _, pos, i = _named_vector(tokens, i, line, 3)
        elif k == "scale":
            _, scale, i = _named_vector(tokens, i, line, 3)
        elif k in ("rotationx", "rotatex"):
            _, rx, i = _named_float(tokens, i, line)
        elif k in ("rotationy", "rotatey"):
            _, ry, i = _named_float(tokens, i, line)
        elif k in ("rotationz", "rotatez"):
            _, rz, i = _named_float(tokens, i, line)
        elif k in ("worldtolocal", "worldtoprimitive", "world-to-local", "matrix"):
            _, w2l, i = _named_matrix4(tokens, i, line)
        elif k in ("smooth", "smoothunion", "smoothunionradius"):
            smooth, i = one(), i + 2
        elif k in ("attr", "attribute", "basecolor", "color", "opacity", "emission", "roughness", "metallic",
                   "transmission"):
            raise SceneScriptError("unsupported", tokens[i], line)
        else:
            raise SceneScriptError("invalidArgumentCount", "sdf", line)
    if shape == "sphere":
        if radius is None:
            raise SceneScriptError("invalidArgumentCount", "sdf", line)
        kw = dict(kind="sphere", radius=radius)
    elif shape == "box":
        if half_ext is None:
            raise SceneScriptError("invalidArgumentCount", "sdf", line)
        kw = dict(kind="box", half_extents=tuple(half_ext), corner_radius=corner)
    else:
        if radius is None or half_height is None:
            raise SceneScriptError("invalidArgumentCount", "sdf", line)
        kw = dict(kind="cylinder", radius=radius, half_height=half_height)
    if mname is None:
        raise SceneScriptError("invalidArgumentCount", "sdf", line)
    mat = _material_id(mname, line, ctx)
    if w2l is not None:
        t = Transform(np.linalg.inv(w2l.astype(np.float64)).astype(f32))
    else:
        t = Transform.translation(pos)
        if rx is not None:
            t = t * Transform.rotation_x(rx)
        if ry is not None:
            t = t * Transform.rotation_y(ry)
        if rz is not None:
            t = t * Transform.rotation_z(rz)
        t = t * Transform.scale(scale)
    return SDFPrimitive(material=mat, transform=t, smooth_union_radius=smooth, operation=op, **kw), i


def _parse_sdf(tokens, line, scene, ctx) -> None:
    """parseSDFVolume (SceneScript.swift:1114-1304), baked with this
    backend's builders at a cubic resolution."""
    from . import sdf as S
    if len(tokens) < 2:
        raise SceneScriptError("invalidArgumentCount", "sdf", line)
    sparse, fallback = False, None
    res = max(2, scene.render_defaults.sdf_resolution or 32)
    brick, sample_scale, band = 8, 1, 0.1
    bmin, bmax = [-1.0, -1.0, -1.0], [1.0, 1.0, 1.0]
    pos, scale, rot = [0.0, 0.0, 0.0], [1.0, 1.0, 1.0], None
    prims = []
    i = 1
    if i < len(tokens) and tokens[i].lower() not in _SDF_GLOBALS and tokens[i].lower() not in _SDF_SHAPES \
            and tokens[i].lower() not in _SDF_OPS:
        i += 1  # the volume's name
    while i < len(tokens):
        k = tokens[i].lower()
        if k in _SDF_SHAPES or k in _SDF_OPS:
            p, i = _parse_sdf_primitive(tokens, i, line, fallback, ctx)
            prims.append(p)
            continue
        if k == "dense":
            sparse, i = False, i + 1
        elif k in ("sparse", "bricks", "brick"):
            sparse, i = True, i + 1
        elif k in ("backend", "storage"):
            if i + 1 >= len(tokens):
                raise SceneScriptError("invalidArgumentCount", "sdf", line)
            v = tokens[i + 1].lower()
            if v == "dense":
                sparse = False
            elif v in ("sparse", "bricks", "brick"):
                sparse = True
            else:
                raise SceneScriptError("invalidArgumentCount", "sdf", line)
            i += 2
        elif k == "material":
            fallback, i = _optional_named_string(tokens, i, "material", False, line)
        elif k in ("resolution", "res", "bricksize", "brick-size", "samplescale", "sample-scale", "sdfsamplescale",
                   "sdf-sample-scale", "narrowband", "band"):
            if i + 1 >= len(tokens):
                raise SceneScriptError("invalidArgumentCount", "sdf", line)
            if k in ("resolution", "res"):
                res = max(2, _int(tokens[i + 1], line))
            elif k in ("bricksize", "brick-size"):
                brick = max(1, _int(tokens[i + 1], line))
            elif k in ("narrowband", "band"):
                band = _float1(tokens, i + 1, line)
            else:
                sample_scale = max(1, _int(tokens[i + 1], line))
            i += 2
        elif k in ("boundsmin", "min"):
            _, bmin, i = _named_vector(tokens, i, line, 3)
        elif k in ("boundsmax", "max"):
            _, bmax, i = _named_vector(tokens, i, line, 3)
        elif k in ("attributes", "attrs"):
            raise SceneScriptError("unsupported", tokens[i], line)
        elif k in ("position", "translation", "translate"):
            _, pos, i = _named_vector(tokens, i, line, 3)
        elif k == "scale":
            _, scale, i = _named_vector(tokens, i, line, 3)
        elif k in ("rotationy", "rotatey"):
            _, rot, i = _named_float(tokens, i, line)
        else:
            raise SceneScriptError("invalidArgumentCount", "sdf", line)
    if not prims:
        raise SceneScriptError("invalidArgumentCount", "sdf", line)
    fallback_mat = _material_id(fallback, line, ctx) if fallback is not None else prims[0].material
    t = Transform.translation(pos)
    if rot is not None:
        t = t * Transform.rotation_y(rot)
    t = t * Transform.scale(scale)
    r = max(2, ctx.sdf_override or res)
    if sparse:
        vol = S.build_sparse_reference(prims, (r, r, r), bmin, bmax, brick_size=brick, narrow_band=band,
                                       sample_scale=sample_scale)
        scene.add_sparse_volume(vol, fallback_mat, t)
    else:
        vol = S.build_volume(prims, (r, r, r), bmin, bmax)
        scene.add_volume(vol, fallback_mat, t)


# ------------------------------------------------------------------ driver --
def _parse_lines(source: str, scene: RenderScene, ctx: _Ctx) -> None:
    for ln, raw in enumerate(source.split("\n")):
        line = ln + 1
        text = _strip_comment(raw).strip()
        if not text:
            continue
        tokens = _tokenize(text)
        if not tokens:
            continue
        cmd = tokens[0].lower()
        if cmd == "include":
            if len(tokens) != 2:
                raise SceneScriptError("invalidArgumentCount", "include", line)
            name = tokens[1]
            if ctx.include_resolver is None:
                raise SceneScriptError("includeResolverMissing", name, line)
            if name in ctx.include_stack:
                raise SceneScriptError("includeCycle", name, line)
            ctx.include_stack.append(name)
            _parse_lines(ctx.include_resolver(name), scene, ctx)
            ctx.include_stack.pop()
        elif cmd in ("render", "defaults", "renderdefaults"):
            scene.render_defaults = _parse_render_defaults(tokens, line, ctx.base_dir, scene.render_defaults)
        elif cmd == "camera":
            scene.camera = _parse_camera(tokens, line)
        elif cmd == "environment":
            scene.environment = _parse_environment(tokens, line, ctx)
        elif cmd == "texture":
            name, tex = _parse_texture(tokens, line, ctx)
            ctx.textures[name] = tex
        elif cmd == "mesh":
            name, mesh = _parse_mesh(tokens, line, ctx)
            ctx.meshes[name] = mesh
        elif cmd == "material":
            name, mat = _parse_material(tokens, line, ctx)
            ctx.materials[name] = scene.add_material(mat)
        elif cmd in ("sdf", "volume"):
            _parse_sdf(tokens, line, scene, ctx)
        elif cmd == "quad":
            mesh, mat = _parse_quad(tokens, line, ctx)
            scene.add(mesh, mat)
        elif cmd == "box":
            mesh, mat, t = _parse_box(tokens, line, ctx)
            scene.add(mesh, mat, t)
        elif cmd == "instance":
            mesh, mat, t = _parse_instance(tokens, line, ctx)
            scene.add(mesh, mat, t)
        else:
            raise SceneScriptError("unknownCommand", cmd, line)


def parse(source: str, base_dir: Optional[str] = None, include_resolver: Optional[IncludeResolver] = None,
          sdf_resolution_override: Optional[int] = None) -> RenderScene:
    """SceneScript.parse(_:baseURL:...) (SceneScript.swift:73-100): relative
    asset paths resolve against `base_dir` (else the working directory);
    ``include`` needs `include_resolver` (name -> source text)."""
    scene = RenderScene()
    scene.render_defaults = RenderDefaults()
    _parse_lines(source, scene, _Ctx(base_dir, include_resolver, sdf_resolution_override))
    return scene


def parse_file(path: str, sdf_resolution_override: Optional[int] = None) -> RenderScene:
    """SceneScript.parse(contentsOf:) (SceneScript.swift:103-123): assets and
    includes resolve beside the file."""
    with open(path, "r", encoding="utf-8") as f:
        source = f.read()
    base = os.path.dirname(os.path.abspath(path))

    def resolver(name: str) -> str:
        with open(_asset_path(name, base), "r", encoding="utf-8") as g:
            return g.read()

    return parse(source, base, resolver, sdf_resolution_override)


def render_file(path: str, out: Optional[str] = None, samples: Optional[int] = None,
                size: Optional[Tuple[int, int]] = None, device: int = 0, sdf_resolution_override: Optional[int] = None):
    """The CLI render of a script (DenrimCLI/main.swift, ``denrim scene.denrim``)
    on this backend: the script's render defaults, overridden by the
    arguments, then the output written as PNG. Returns the PNG path."""
    from .renderer import DenrimRenderer
    scene = parse_file(path, sdf_resolution_override)
    d = scene.render_defaults
    settings = d.render_settings()
    if size is not None:
        import dataclasses
        settings = dataclasses.replace(settings, width=size[0], height=size[1])
    n = samples if samples is not None else (d.samples if d.samples is not None else 64)
    target = out or d.output_path or os.path.splitext(path)[0] + ".png"
    r = DenrimRenderer(device)
    gs = r.upload(compile_scene(scene))
    sess = r.make_session(gs, settings, camera=scene.camera)
    sess.render(n)
    sess.write_png(target, d.output if d.output is not None else RenderOutput.BEAUTY)
    sess.close()
    gs.release()
    return target


def main(argv=None) -> None:
    import argparse
    ap = argparse.ArgumentParser(prog="python -m denrimrendererkit_amd.scenescript",
                                 description="Render a .denrim scene script on the HIP backend.")
    ap.add_argument("script")
    ap.add_argument("--out")
    ap.add_argument("--spp", type=int)
    ap.add_argument("--size", help="WxH")
    ap.add_argument("--sdf-resolution", type=int)
    a = ap.parse_args(argv)
    size = tuple(int(v) for v in a.size.split("x")) if a.size else None
    print(render_file(a.script, a.out, a.spp, size, sdf_resolution_override=a.sdf_resolution))


if __name__ == "__main__":
    main()
