"""Built-in material presets (Scene/BuiltInMaterialLibrary.swift:59-321).

``PRESETS`` lists (identifier, category, Material) in the reference's order
with the reference's parameter values (the table is data: it has to match
for ``material X preset metal.gold`` to mean the same material);
``material(name)`` looks one up case-insensitively with ``_`` read as ``-``
(:398-409) and returns a copy, or None.
"""
from __future__ import annotations

import dataclasses
from typing import List, Optional, Tuple

from .scene import Material

M = Material

PRESETS: List[Tuple[str, str, Material]] = [
    ("matte.clay", "diffuse", M(base_color=(0.78, 0.62, 0.48), roughness=0.82, specular=0.25)),
    ("matte.paper", "diffuse", M(base_color=(0.92, 0.9, 0.86), roughness=0.9, specular=0.18)),
    ("matte.charcoal", "diffuse", M(base_color=(0.025, 0.025, 0.023), roughness=0.86, specular=0.12)),
    ("plastic.soft-black", "plastic", M(base_color=(0.035, 0.034, 0.032), roughness=0.38, specular=0.55,
                                        clearcoat=0.18, clearcoat_roughness=0.24)),
    ("plastic.gloss-white", "plastic", M(base_color=(0.86, 0.86, 0.82), roughness=0.18, specular=0.7,
                                         clearcoat=0.45, clearcoat_roughness=0.08)),
    ("rubber.black", "plastic", M(base_color=(0.012, 0.012, 0.011), roughness=0.78, specular=0.22)),
    ("fabric.velvet-purple", "fabric", M(base_color=(0.22, 0.08, 0.42), roughness=0.74, specular=0.18, sheen=0.72,
                                         sheen_color=(0.72, 0.52, 1.0), sheen_roughness=0.82)),
    ("fabric.denim-blue", "fabric", M(base_color=(0.055, 0.12, 0.26), roughness=0.88, specular=0.16, sheen=0.3,
                                      sheen_color=(0.35, 0.5, 0.78), sheen_roughness=0.9)),
    ("organic.wood", "diffuse", M(base_color=(0.48, 0.25, 0.105), roughness=0.78, specular=0.24, sheen=0.12,
                                  sheen_color=(0.82, 0.58, 0.32), sheen_roughness=0.86)),
    ("organic.plant", "subsurface", M(base_color=(0.14, 0.42, 0.105), roughness=0.66, specular=0.28, sheen=0.18,
                                      sheen_color=(0.48, 0.86, 0.28), sheen_roughness=0.72, subsurface=0.26,
                                      subsurface_color=(0.34, 0.82, 0.18), subsurface_radius=(0.32, 0.78, 0.22),
                                      subsurface_scale=0.18, subsurface_anisotropy=0.08)),
    ("subsurface.skin-warm", "subsurface", M(base_color=(0.82, 0.52, 0.38), roughness=0.56, specular=0.38,
                                             index_of_refraction=1.38, subsurface=0.82,
                                             subsurface_color=(0.95, 0.48, 0.32), subsurface_radius=(1.0, 0.42, 0.24),
                                             subsurface_scale=0.34, subsurface_anisotropy=0.18)),
    ("subsurface.wax-cream", "subsurface", M(base_color=(0.94, 0.82, 0.58), roughness=0.48, specular=0.42,
                                             index_of_refraction=1.44, subsurface=0.88,
                                             subsurface_color=(0.98, 0.82, 0.48), subsurface_radius=(1.3, 0.88, 0.42),
                                             subsurface_scale=0.48, subsurface_anisotropy=0.0)),
    ("liquid.milk", "liquid", M(base_color=(0.98, 0.96, 0.88), roughness=0.68, specular=0.18,
                                index_of_refraction=1.35, transmission=0.46, transmission_color=(0.99, 0.985, 0.94),
                                transmission_roughness=0.58, transmission_index_of_refraction=1.35,
                                transmission_absorption_color=(0.7, 0.72, 0.58),
                                transmission_absorption_distance=0.62, volume_scattering=0.76,
                                volume_scattering_color=(0.98, 0.97, 0.9), volume_scattering_distance=0.5,
                                volume_anisotropy=0.25)),
    ("subsurface.jade-green", "subsurface", M(base_color=(0.28, 0.62, 0.48), roughness=0.42, specular=0.48,
                                              index_of_refraction=1.56, clearcoat=0.14, clearcoat_roughness=0.16,
                                              subsurface=0.7, subsurface_color=(0.38, 0.86, 0.58),
                                              subsurface_radius=(0.38, 0.9, 0.52), subsurface_scale=0.24,
                                              subsurface_anisotropy=0.08)),
    ("metal.chrome", "metal", M(base_color=(0.95, 0.95, 0.93), roughness=0.012, metallic=1.0)),
    ("metal.aluminum", "metal", M(base_color=(0.84, 0.85, 0.84), roughness=0.08, metallic=1.0)),
    ("metal.brushed-aluminum", "metal", M(base_color=(0.84, 0.85, 0.84), roughness=0.16, metallic=1.0,
                                          specular_anisotropy=0.72)),
    ("metal.gold", "metal", M(base_color=(1.0, 0.72, 0.29), roughness=0.06, metallic=1.0)),
    ("metal.copper", "metal", M(base_color=(0.98, 0.45, 0.24), roughness=0.07, metallic=1.0)),
    ("coating.car-red", "coating", M(base_color=(0.72, 0.025, 0.018), roughness=0.34, specular=0.55, clearcoat=0.9,
                                     clearcoat_color=(1.0, 0.92, 0.88), clearcoat_roughness=0.045)),
    ("coating.pearl-white", "coating", M(base_color=(0.86, 0.84, 0.78), roughness=0.28, specular=0.6, clearcoat=0.85,
                                         clearcoat_color=(1.0, 0.94, 0.82), clearcoat_roughness=0.04)),
    ("coating.iridescent-amber", "coating", M(base_color=(0.88, 0.24, 0.035), roughness=0.18, specular=0.72,
                                              specular_color=(1.0, 0.9, 0.74), clearcoat=0.95,
                                              clearcoat_color=(1.0, 0.68, 0.2), clearcoat_roughness=0.035,
                                              clearcoat_index_of_refraction=1.55, thin_film=0.48,
                                              thin_film_thickness_nanometers=520.0,
                                              thin_film_index_of_refraction=1.38)),
    ("glass.clear", "glass", M(base_color=(0.92, 0.98, 0.98), roughness=0.01, specular=1.0, index_of_refraction=1.45,
                               transmission=1.0, transmission_color=(0.96, 1.0, 1.0), transmission_roughness=0.01,
                               transmission_index_of_refraction=1.45)),
    ("glass.frosted", "glass", M(base_color=(0.75, 0.9, 0.95), roughness=0.32, specular=0.9, index_of_refraction=1.45,
                                 transmission=0.92, transmission_color=(0.82, 0.94, 1.0), transmission_roughness=0.3,
                                 transmission_index_of_refraction=1.45)),
    ("glass.thin-pane", "glass", M(base_color=(0.9, 0.98, 1.0), roughness=0.012, specular=1.0,
                                   index_of_refraction=1.45, transmission=1.0, transmission_color=(0.94, 1.0, 1.0),
                                   transmission_roughness=0.012, transmission_index_of_refraction=1.45,
                                   thin_walled=True)),
    ("liquid.water", "liquid", M(base_color=(0.72, 0.9, 1.0), roughness=0.002, specular=1.0,
                                 index_of_refraction=1.333, transmission=1.0, transmission_color=(0.8, 0.95, 1.0),
                                 transmission_roughness=0.002, transmission_index_of_refraction=1.333)),
    ("ceramic.white", "ceramic", M(base_color=(0.82, 0.8, 0.76), roughness=0.18, specular=0.65, clearcoat=0.58,
                                   clearcoat_roughness=0.12)),
    ("ceramic.porcelain", "ceramic", M(base_color=(0.9, 0.88, 0.82), roughness=0.1, specular=0.72, clearcoat=0.5,
                                       clearcoat_roughness=0.07)),
    ("emission.warm-panel", "emission", M(base_color=(1.0, 0.92, 0.8), emission=(1.0, 0.78, 0.55),
                                          emission_strength=4.0, roughness=0.5, specular=0.0)),
    ("emission.cool-panel", "emission", M(base_color=(0.82, 0.9, 1.0), emission=(0.65, 0.78, 1.0),
                                          emission_strength=4.0, roughness=0.5, specular=0.0)),
]


def _normalized(name: str) -> str:
    return name.strip().lower().replace("_", "-")


IDENTIFIERS = [p[0] for p in PRESETS]


def material(name: str) -> Optional[Material]:
    """BuiltInMaterialLibrary.material(named:): a copy of the preset, or None."""
    key = _normalized(name)
    for ident, _, m in PRESETS:
        if _normalized(ident) == key:
            return dataclasses.replace(m)
    return None
