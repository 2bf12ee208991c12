"""OBJ ingestion (sr_obj_parse), host side: the loader against an
independent reading of the format as tobj 0.1.3 defines it (the crate the
reference's examples use, examples/suzanne.rs:23-47): distinct v/vt/vn
corners become vertices in order of first use, polygons become fans from
their first corner, `o` starts a model."""
import os

import numpy as np
import pytest

import softrender_hip as sr

HERE = os.path.dirname(os.path.abspath(__file__))
SUZANNE = os.path.join(HERE, "golden", "suzanne.obj")


def read_obj(text):
    """Independent reading: positions/uvs/normals and faces per model."""
    P, T, N, models, faces = [], [], [], [], []
    for line in text.splitlines():
        t = line.split()
        if not t:
            continue
        if t[0] == "v":
            P.append([float(x) for x in t[1:4]])
        elif t[0] == "vt":
            T.append([float(x) for x in t[1:3]])
        elif t[0] == "vn":
            N.append([float(x) for x in t[1:4]])
        elif t[0] == "f":
            face = []
            for c in t[1:]:
                parts = (c.split("/") + ["", ""])[:3]
                ids = []
                for k, (p, arr) in enumerate(zip(parts, (P, T, N))):
                    if p == "":
                        ids.append(None)
                    else:
                        i = int(p)
                        ids.append(i - 1 if i > 0 else len(arr) + i)
                face.append(tuple(ids))
            faces.append(face)
        elif t[0] == "o" and faces:
            models.append(faces)
            faces = []
    if faces:
        models.append(faces)
    return np.array(P, np.float32), np.array(T, np.float32), np.array(N, np.float32), models


def expected_model(P, T, N, faces):
    corners, order = {}, []
    idx = []
    for face in faces:
        for k in range(2, len(face)):
            for c in (face[0], face[k - 1], face[k]):
                if c not in corners:
                    corners[c] = len(order)
                    order.append(c)
                idx.append(corners[c])
    pos = np.array([P[c[0]] for c in order], np.float32)
    uv = np.array([T[c[1]] for c in order], np.float32) if order and order[0][1] is not None else None
    nrm = np.array([N[c[2]] for c in order], np.float32) if order and order[0][2] is not None else None
    return pos, uv, nrm, np.array(idx, np.uint32)


def test_suzanne_matches_the_format():
    P, T, N, models = read_obj(open(SUZANNE).read())
    got = sr.parse_obj(SUZANNE)
    assert got["models"] == len(models) == 1
    faces = models[0]
    assert sum(len(f) - 2 for f in faces) * 3 == len(got["indices"]) == (32 + 2 * 468) * 3
    pos, uv, nrm, idx = expected_model(P, T, N, faces)
    assert np.array_equal(got["indices"], idx)
    assert np.array_equal(got["positions"], pos)
    assert np.array_equal(got["normals"], nrm)
    assert np.array_equal(got["uvs"], uv)
    # vertices are numbered in order of first use
    first = np.maximum.accumulate(np.concatenate([[-1], got["indices"][:-1].astype(np.int64)]))
    assert np.all(got["indices"] <= first + 1)
    # quads split along the diagonal from their first corner: triangle 2q+1
    # of a quad (a, b, c, d) is (a, c, d)
    quad = next(i for i, f in enumerate(faces) if len(f) == 4)
    t0 = sum(len(f) - 2 for f in faces[:quad])
    a, b, c, d = faces[quad]
    tri = got["positions"][got["indices"][3 * (t0 + 1):3 * (t0 + 2)]]
    assert np.array_equal(tri, P[[a[0], c[0], d[0]]])


def test_indices_forms_fans_and_models(tmp_path):
    text = """# comment
mtllib x.mtl
o first
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 0.5 1.5 0
vt 0 0
vt 1 1
vn 0 0 1
vn 0 0 -1
f 1//1 2//1 3//1 5//1 4//1
f -5//-2 -4//-2 -3//-1
o second
v 2 2 2
f -1/1/1 -2/2/1 -3/1/2
"""
    p = tmp_path / "t.obj"
    p.write_text(text)
    P, T, N, models = read_obj(text)
    assert len(models) == 2
    for m in range(2):
        got = sr.parse_obj(str(p), m)
        pos, uv, nrm, idx = expected_model(P, T, N, models[m])
        assert got["models"] == 2
        assert np.array_equal(got["indices"], idx)
        assert np.array_equal(got["positions"], pos)
        assert np.array_equal(got["normals"], nrm)
        if uv is None:
            assert got["uvs"] is None
        else:
            assert np.array_equal(got["uvs"], uv)
    assert len(sr.parse_obj(str(p), 0)["indices"]) == 3 * (3 + 1)  # pentagon fan + triangle


def test_errors(tmp_path):
    with pytest.raises(sr.SoftrenderError):
        sr.parse_obj(str(tmp_path / "missing.obj"))
    bad = tmp_path / "bad.obj"
    bad.write_text("v 0 0 0\nv 1 0 0\nf 1 2 7\n")
    with pytest.raises(sr.SoftrenderError):
        sr.parse_obj(str(bad))
