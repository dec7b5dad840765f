"""Writes small glTF 2.0 files for the loader tests, straight from the specification
(GLB container: 12-byte header, JSON chunk padded with spaces, BIN chunk padded with zeros;
matrices column-major; accessors over bufferViews with optional byteStride)."""
from __future__ import annotations

import base64
import json
import struct

import numpy as np

FLOAT, U8, U16, U32 = 5126, 5121, 5123, 5125


class Builder:
    def __init__(self):
        self.bin = bytearray()
        self.doc = {"asset": {"version": "2.0"}, "buffers": [], "bufferViews": [], "accessors": []}

    def view(self, data: bytes, stride: int | None = None) -> int:
        while len(self.bin) % 4:
            self.bin += b"\0"
        bv = {"buffer": 0, "byteOffset": len(self.bin), "byteLength": len(data)}
        if stride:
            bv["byteStride"] = stride
        self.bin += data
        self.doc["bufferViews"].append(bv)
        return len(self.doc["bufferViews"]) - 1

    def accessor(self, view: int, comp: int, typ: str, count: int, offset: int = 0, normalized: bool = False) -> int:
        a = {"bufferView": view, "componentType": comp, "type": typ, "count": count}
        if offset:
            a["byteOffset"] = offset
        if normalized:
            a["normalized"] = True
        self.doc["accessors"].append(a)
        return len(self.doc["accessors"]) - 1

    def finish(self) -> tuple[dict, bytes]:
        while len(self.bin) % 4:
            self.bin += b"\0"
        self.doc["buffers"] = [{"byteLength": len(self.bin)}]
        return self.doc, bytes(self.bin)


def glb_bytes(doc: dict, bin_: bytes | None) -> bytes:
    js = json.dumps(doc).encode()
    js += b" " * (-len(js) % 4)
    out = struct.pack("<II", len(js), 0x4E4F534A) + js
    if bin_ is not None:
        b = bin_ + b"\0" * (-len(bin_) % 4)
        out += struct.pack("<II", len(b), 0x004E4942) + b
    return struct.pack("<III", 0x46546C67, 2, 12 + len(out)) + out


def write_glb(path, doc: dict, bin_: bytes | None) -> None:
    with open(path, "wb") as f:
        f.write(glb_bytes(doc, bin_))


def write_gltf_data_uri(path, doc: dict, bin_: bytes) -> None:
    d = json.loads(json.dumps(doc))
    d["buffers"][0]["uri"] = "data:application/octet-stream;base64," + base64.b64encode(bin_).decode()
    with open(path, "w") as f:
        json.dump(d, f, indent=1)


def write_gltf_external(path, doc: dict, bin_: bytes, bin_name: str) -> None:
    """bin_name is written percent-encoded (spaces -> %20) as the buffer's uri."""
    import os

    d = json.loads(json.dumps(doc))
    d["buffers"][0]["uri"] = bin_name.replace(" ", "%20")
    with open(os.path.join(os.path.dirname(str(path)), bin_name), "wb") as f:
        f.write(bin_)
    with open(path, "w") as f:
        json.dump(d, f)


def png_header(w: int, h: int) -> bytes:
    """The first 33 bytes of a PNG: signature and IHDR (enough for the loader's size probe)."""
    ihdr = struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)
    return b"\x89PNG\r\n\x1a\n" + struct.pack(">I", 13) + b"IHDR" + ihdr + b"\0\0\0\0"


def scene_to_gltf(sa) -> tuple[dict, bytes]:
    """A SceneArrays as glTF: one mesh (one primitive) per rrh mesh, its vertices as one
    interleaved bufferView (stride 48: pos, normal, uv, tangent), u32 indices; one material per
    record (factors only); one root node per instance carrying its matrix (column-major)."""
    b = Builder()
    meshes = []
    for m in sa.meshes:
        vb, vc = int(m["vertex_base"]), int(m["vertex_count"])
        v = b.view(sa.verts[vb:vb + vc].tobytes(), stride=48)
        attrs = {"POSITION": b.accessor(v, FLOAT, "VEC3", vc, 0), "NORMAL": b.accessor(v, FLOAT, "VEC3", vc, 12),
                 "TEXCOORD_0": b.accessor(v, FLOAT, "VEC2", vc, 24), "TANGENT": b.accessor(v, FLOAT, "VEC4", vc, 32)}
        i0, ic = int(m["index_byte_offset"]) // 4, int(m["index_count"])
        iv = b.view(sa.idx[i0:i0 + ic].astype("<u4").tobytes())
        meshes.append({"primitives": [{"attributes": attrs, "indices": b.accessor(iv, U32, "SCALAR", ic),
                                       "material": int(m["material_index"])}]})
    mats = []
    for mt in sa.mats:
        mats.append({"pbrMetallicRoughness": {"baseColorFactor": [float(x) for x in mt["base_color"]],
                                              "metallicFactor": float(mt["metallic"]),
                                              "roughnessFactor": float(mt["roughness"])},
                     "emissiveFactor": [float(x) for x in mt["emissive"]]})
    nodes = []
    for inst in sa.insts:
        t = np.asarray(inst["transform"], np.float64).reshape(4, 4)  # row-major
        nodes.append({"mesh": int(inst["mesh_index"]), "matrix": [float(x) for x in t.T.reshape(-1)]})
    doc, bin_ = b.finish()
    doc.update(meshes=meshes, materials=mats, nodes=nodes, scenes=[{"nodes": list(range(len(nodes)))}], scene=0)
    return doc, bin_
