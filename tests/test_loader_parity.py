"""libsbr's model loader against the reference's own loader.

The reference loader (src/modelLoader.cpp with the tinyobjloader and
tinygltf it vendors) is compiled from the reference sources into
oracle/_ref/libref_loader.so by oracle/Makefile. Positions and indices must
be bit-identical:

  * on every OBJ model the reference ships (assets/test_models/*.obj);
  * on generated OBJ files with triangles, quads, convex and concave
    polygons of 5+ corners (tinyobj's ear clipping), v/vt/vn corners that
    split vertices, and relative indices;
  * on generated glTF (data-URI and external buffers) and GLB files with
    u8 / u16 / u32 / implicit indices and several primitives.

Both sides build the quad split and the ear clipping without FMA contraction
(the reference's own Release build, with -march=native, may let the compiler
contract them and split a near-square quad the other way).
"""
import base64
import glob
import json
import os
import struct

import numpy as np
import pytest

import oracle_lib as O

S = pytest.importorskip("sagittasbr_amd")
from sagittasbr_amd import meshes as M  # noqa: E402

REF_MODELS = sorted(glob.glob("/root/reference/assets/test_models/*.obj"))


def _need():
    if O.ref_loader() is None:
        pytest.skip("reference loader not built (oracle/Makefile builds it where the reference tree exists)")


def _same(path):
    rv, rf = O.ref_load_model(path)
    v, f = S.load_model(path)
    assert v.shape == rv.shape and f.shape == rf.shape, (v.shape, rv.shape, f.shape, rf.shape)
    assert np.array_equal(v.view(np.uint32), rv.view(np.uint32))
    assert np.array_equal(f, rf)
    return v, f


@pytest.mark.parametrize("path", REF_MODELS, ids=[os.path.basename(p) for p in REF_MODELS])
def test_reference_obj_models(path):
    _need()
    v, f = _same(path)
    assert len(f) > 0


def test_generated_obj_polygons(tmp_path):
    _need()
    rng = np.random.default_rng(5)
    lines = []
    nv = 0

    def add_v(p):
        nonlocal nv
        lines.append("v %.9g %.9g %.9g" % tuple(p))
        nv += 1
        return nv

    # convex n-gons in tilted planes, n = 3..9
    for n in range(3, 10):
        c = rng.normal(size=3)
        a, b = np.linalg.qr(rng.normal(size=(3, 3)))[0][:, :2].T
        ids = [add_v(c + np.cos(2 * np.pi * k / n) * a + np.sin(2 * np.pi * k / n) * b) for k in range(n)]
        lines.append("f " + " ".join(str(i) for i in ids))
    # concave polygons: an L, an arrow, a star, each in a coordinate plane
    shapes = [
        [(0, 0), (3, 0), (3, 1), (1, 1), (1, 3), (0, 3)],
        [(0, 0), (2, 1), (4, 0), (2, 3)],
        [(np.cos(k * np.pi / 5) * (1.0 if k % 2 == 0 else 0.4), np.sin(k * np.pi / 5) * (1.0 if k % 2 == 0 else 0.4))
         for k in range(10)],
    ]
    for s, pts in enumerate(shapes):
        for plane in range(3):
            ids = []
            for x, y in pts:
                p = [0.25 * s, 0.5, 0.75]
                p[plane], p[(plane + 1) % 3] = x + s, y - s
                ids.append(add_v(p))
            lines.append("f " + " ".join(str(i) for i in ids))
            lines.append("f " + " ".join(str(i) for i in reversed(ids)))
    # corners with texcoords and normals (vertex splits) and relative indices
    lines += ["vt 0 0", "vt 1 0", "vt 1 1", "vn 0 0 1", "vn 0 1 0"]
    lines.append("f -1/1/1 -2/2/1 -3/3/2 -4/1/2 -5//1")
    lines.append("f 1/1 2/2 3/3")
    lines.append("f 1//2 2//2 3//1 4//1")
    p = tmp_path / "poly.obj"
    p.write_text("\n".join(lines) + "\n")
    v, f = _same(str(p))
    assert len(f) > 60


def _gltf(prims, uri_kind):
    """glTF document + binary blob holding every primitive's positions and
    indices; prims = [(verts, faces, index componentType or None)]."""
    blob = b""
    views, accs, out = [], [], []
    for verts, faces, ctype in prims:
        att = {}
        vb = np.asarray(verts, np.float32).tobytes()
        views.append({"buffer": 0, "byteOffset": len(blob), "byteLength": len(vb)})
        blob += vb
        accs.append({"bufferView": len(views) - 1, "componentType": 5126, "count": len(verts), "type": "VEC3"})
        att["POSITION"] = len(accs) - 1
        prim = {"attributes": att, "mode": 4}
        if ctype is not None:
            dt = {5121: np.uint8, 5123: np.uint16, 5125: np.uint32}[ctype]
            ib = np.asarray(faces, dt).reshape(-1).tobytes()
            ib += b"\0" * ((4 - len(ib) % 4) % 4)
            views.append({"buffer": 0, "byteOffset": len(blob), "byteLength": int(np.size(faces)) * np.dtype(dt).itemsize})
            blob += ib
            accs.append({"bufferView": len(views) - 1, "componentType": ctype, "count": int(np.size(faces)),
                         "type": "SCALAR"})
            prim["indices"] = len(accs) - 1
        out.append(prim)
    doc = {"asset": {"version": "2.0"}, "buffers": [{"byteLength": len(blob)}], "bufferViews": views,
           "accessors": accs, "meshes": [{"primitives": out[:2]}, {"primitives": out[2:]}]}
    if uri_kind == "data":
        doc["buffers"][0]["uri"] = "data:application/octet-stream;base64," + base64.b64encode(blob).decode()
    elif uri_kind == "file":
        doc["buffers"][0]["uri"] = "mesh.bin"
    return doc, blob


@pytest.mark.parametrize("kind", ["data", "file", "glb"])
def test_generated_gltf(tmp_path, kind):
    _need()
    v1, f1 = M.icosphere(1)
    v2, f2 = M.icosphere(2)
    v3, f3 = M.trihedral_duct_scene()
    small = (v1[:12], f1[:20] % 12)
    prims = [(small[0], small[1], 5121), (v2, f2, 5123), (v3, f3, 5125),
             (v1[:9], np.arange(9).reshape(3, 3), None)]
    doc, blob = _gltf(prims, "glb" if kind == "glb" else kind)
    if kind == "glb":
        js = json.dumps(doc).encode()
        js += b" " * ((4 - len(js) % 4) % 4)
        bb = blob + b"\0" * ((4 - len(blob) % 4) % 4)
        data = struct.pack("<4sII", b"glTF", 2, 12 + 8 + len(js) + 8 + len(bb))
        data += struct.pack("<II", len(js), 0x4E4F534A) + js + struct.pack("<II", len(bb), 0x004E4942) + bb
        p = tmp_path / "mesh.glb"
        p.write_bytes(data)
    else:
        if kind == "file":
            (tmp_path / "mesh.bin").write_bytes(blob)
        p = tmp_path / "mesh.gltf"
        p.write_text(json.dumps(doc))
    v, f = _same(str(p))
    assert len(f) == 20 + len(f2) + len(f3) + 3
