"""OBJ input and frame export (host C++, lats_obj_* in the C ABI).

The reader is checked on hand-written files whose triangles follow from the
OBJ rules the reference's reader applies (tinyobjloader defaults: fan
triangulation from the first corner, 1-based and negative indices, v/vt/vn
corners; mesh_reader.cpp:172-182 drops a triangle that repeats a vertex), and
by a round trip of the square8K mesh (the reference's resources/models/
square8K.obj, fixture tests/golden/square8k_mesh.npz) through the writer,
whose colouring must then give the reference's published 10 clusters.
"""
import numpy as np
import pytest

from libatssim_amd import _lib
from libatssim_amd.vbd import Topology, parse_obj, read_obj, write_obj
from tests.helpers import square8k_mesh


def test_fan_triangulation_corner_forms_and_negative_indices():
    text = """# a quad, a pentagon and relative indices
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
vt 0 0
vn 0 0 1
f 1/1/1 2/1/1 3/1/1 4/1/1
v 2 0 0
v 2 1 0 1.0
f 2//1 5//1 6//1 3//1 -3
f -2 -1 -4
"""
    x, f, dropped = parse_obj(text)
    assert x.shape == (6, 3) and x[5].tolist() == [2.0, 1.0, 0.0]
    assert f.tolist() == [[0, 1, 2], [0, 2, 3],          # quad 1 2 3 4: fan from its first corner
                          [1, 4, 5], [1, 5, 2], [1, 2, 3],   # pentagon 2 5 6 3 -3 (-3 = the 4th vertex)
                          [4, 5, 2]]                     # -2 -1 -4 after six vertices = 5 6 3
    assert dropped == 0


def test_degenerate_triangles_are_dropped():
    x, f, dropped = parse_obj("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 2\nf 1 2 3\nf 3 3 3\n")
    assert f.tolist() == [[0, 1, 2]] and dropped == 2


@pytest.mark.parametrize("text", [
    "v 0 0 0\nv 1 0 0\nf 1 2\n",                 # fewer than three corners
    "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n",      # beyond the vertices
    "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 -4\n",     # relative index before the first vertex
    "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 x\n",      # not a number
    "v 0 0\n",                                   # two coordinates
])
def test_malformed_files_are_rejected(text):
    with pytest.raises(_lib.LatsError):
        parse_obj(text)


def test_square8k_round_trip_and_colouring(tmp_path):
    x, f = square8k_mesh()
    path = tmp_path / "frame_0.obj"
    write_obj(path, x, f)
    lines = path.read_text().split("\n")
    assert lines[0].startswith("#") and lines[1] == "o mesh_0"
    assert sum(l.startswith("v ") for l in lines) == 8282 and sum(l.startswith("f ") for l in lines) == 16154
    x2, f2, dropped = read_obj(path)
    assert dropped == 0 and np.array_equal(f2, f)
    # the stream's default precision: 6 significant digits
    err = np.abs(x2.astype(np.float64) - x) / np.maximum(np.abs(x), 1e-30)
    assert err[np.abs(x) > 1e-6].max() < 5e-6
    t = Topology(x2, f2)
    sizes = [len(c) for c in t.clusters()]
    assert sizes == [1384, 1226, 1194, 1144, 1087, 971, 757, 430, 85, 4]


def test_write_rejects_bad_faces(tmp_path):
    with pytest.raises(_lib.LatsError):
        write_obj(tmp_path / "bad.obj", np.zeros((3, 3), np.float32), np.array([[0, 1, 3]], np.int32))
