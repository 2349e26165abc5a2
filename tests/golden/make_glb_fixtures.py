# SPDX-License-Identifier: MIT
"""Writes tests/golden/glb/<case>.json and <case>.npz from the GLB scenes of
the reference's integrated rigid conformance cases
(conformance/v1/cases/<case>.json names the asset under examples/assets).

Data only: the .json keeps, for every scene-root node that carries
parallel-mater extras (pm_system), its name, its extras, its TRS and its
mesh; the .npz keeps every primitive of those meshes as float32 positions
(POSITION, read as cgltf_accessor_read_float does) and uint32 triangle
indices.  How the loader turns these into bodies, joints and shapes is
restated by the tests, not here.

Run in a checkout that holds the reference:
    python tests/golden/make_glb_fixtures.py /root/reference
"""
import json
import pathlib
import struct
import sys

import numpy as np

CASES = ("constraint-point", "constraint-fixed", "constraint-hinge", "constraint-slider", "constraint-piston",
         "constraint-generic", "constraint-generic-spring", "constraint-motor", "passive-active")
COMPONENT = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5125: np.uint32, 5126: np.float32}
WIDTH = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4}


def read_glb(path):
    data = path.read_bytes()
    magic, version, length = struct.unpack_from("<III", data, 0)
    assert magic == 0x46546C67 and version == 2 and length == len(data), path
    offset, doc, blob = 12, None, None
    while offset < length:
        size, kind = struct.unpack_from("<II", data, offset)
        chunk = data[offset + 8:offset + 8 + size]
        if kind == 0x4E4F534A:
            doc = json.loads(chunk)
        elif kind == 0x004E4942:
            blob = chunk
        offset += 8 + size
    return doc, blob


def accessor(doc, blob, index):
    acc = doc["accessors"][index]
    view = doc["bufferViews"][acc["bufferView"]]
    dtype = np.dtype(COMPONENT[acc["componentType"]])
    width = WIDTH[acc["type"]]
    start = view.get("byteOffset", 0) + acc.get("byteOffset", 0)
    stride = view.get("byteStride", dtype.itemsize * width)
    assert not acc.get("sparse") and not acc.get("normalized", False)
    rows = [np.frombuffer(blob, dtype=dtype, count=width, offset=start + k * stride) for k in range(acc["count"])]
    return np.array(rows, dtype=dtype).reshape(acc["count"], width)


def main(root):
    root = pathlib.Path(root)
    out_dir = pathlib.Path(__file__).resolve().parent / "glb"
    out_dir.mkdir(exist_ok=True)
    for name in CASES:
        case = json.loads((root / "conformance" / "v1" / "cases" / f"{name}.json").read_text())
        asset = case["resources"][0]["definition"]["path"]
        doc, blob = read_glb(root / asset)
        parents = {c for n in doc["nodes"] for c in n.get("children", [])}
        nodes, meshes, arrays = [], {}, {}
        for k, node in enumerate(doc["nodes"]):
            extras = node.get("extras") or {}
            if "pm_system" not in extras and "pm_schema" not in extras:
                continue
            entry = {"index": k, "name": node.get("name"), "extras": extras,
                     "translation": node.get("translation"), "rotation": node.get("rotation"),
                     "scale": node.get("scale"), "mesh": node.get("mesh"),
                     "has_matrix": "matrix" in node, "has_parent": k in parents}
            nodes.append(entry)
            m = node.get("mesh")
            if m is None or m in meshes:
                continue
            prims = []
            for p, prim in enumerate(doc["meshes"][m]["primitives"]):
                mode = prim.get("mode", 4)
                pos = accessor(doc, blob, prim["attributes"]["POSITION"]).astype(np.float32)
                idx = (accessor(doc, blob, prim["indices"]).reshape(-1).astype(np.uint32)
                       if "indices" in prim else np.arange(len(pos), dtype=np.uint32))
                arrays[f"mesh{m}_prim{p}_positions"] = pos
                arrays[f"mesh{m}_prim{p}_indices"] = idx
                prims.append({"mode": mode, "vertices": len(pos), "indices": len(idx)})
            meshes[m] = prims
        fixture = {"source": {"case": f"conformance/v1/cases/{name}.json", "asset": asset},
                   "nodes": nodes, "meshes": {str(k): v for k, v in sorted(meshes.items())}}
        (out_dir / f"{name}.json").write_text(json.dumps(fixture, indent=1) + "\n")
        np.savez_compressed(out_dir / f"{name}.npz", **arrays)
        print(f"wrote {name}: {len(nodes)} nodes, {len(meshes)} meshes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
