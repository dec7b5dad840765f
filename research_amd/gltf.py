"""glTF 2.0 import on the Python side: the records the host layer's loader parses
(GltfLoader::load, src/scene/GltfLoader.cpp:79-363) as numpy arrays in the rrh_* layouts,
packed as SceneArrays ready for core.Scene.

A glTF file carries no analytic lights and no scene flags (the reference's loader adds
neither), so `lights` is empty and `flags` 0; the suggested camera is the default camera
(position (0, 0, 3), looking down -Z, 45 degrees), which the reference restores after a load.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .core import SceneArrays


@dataclass
class GltfInfo:
    name: str
    mesh_names: list
    material_names: list
    material_images: list  # (base_color, metallic_roughness, normal, emissive) image index, -1 = none
    images: list
    warnings: list


def load(path: str) -> tuple[SceneArrays, GltfInfo]:
    """Parse a .glb / .gltf file (the host layer's loader); raises RuntimeError on malformed input."""
    from . import _app  # the pybind11 host module (built by research_amd.build.build_host)

    d = _app.load_gltf(str(path))
    arr = lambda key, dt: np.frombuffer(d[key], dtype=dt).copy()
    sa = SceneArrays(arr("vertices", L.VERTEX), arr("indices", np.uint32), arr("meshes", L.MESH),
                     arr("materials", L.MATERIAL), arr("instances", L.INSTANCE), np.zeros(0, dtype=L.LIGHT), 0,
                     (0.0, 0.0, 3.0), (0.0, 0.0, 0.0), 45.0)
    info = GltfInfo(d["name"], list(d["mesh_names"]), list(d["material_names"]),
                    [tuple(t) for t in d["material_images"]], [dict(i) for i in d["images"]], list(d["warnings"]))
    return sa, info
