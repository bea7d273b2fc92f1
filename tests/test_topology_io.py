"""The topology on disk (lats_topology_save / lats_topology_load): one raw
array per file as the reference's save_to_binary writes them, packed-CSR
adjacency and clusters; the square8K clusters checked id for id against a
committed fixture in that format, which is itself pinned to a brute-force
set-based first-fit colouring and the published cluster sizes."""
import ctypes
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import GOLDEN, square8k_mesh
from tests.test_oracle_pinning import brute_edges, first_fit
from libatssim_amd import Topology, _lib, configs
from libatssim_amd._lib import LatsError, check, lib

FILES = ["faces", "edges", "bending_edges", "edges_rest_state_length", "bending_edges_Q", "vert_adj_edges",
         "vert_adj_bending_edges", "clusterd_per_vertex_bending", "per_vertex_bending_cluster_id"]


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _save_with_library(x, f, directory):
    x = np.ascontiguousarray(x, np.float32)
    f = np.ascontiguousarray(f, np.int32)
    h = ctypes.c_void_p()
    check(lib().lats_topology_build(_p(x), len(x), _p(f), len(f), ctypes.byref(h)))
    try:
        check(lib().lats_topology_save(h, os.fsencode(str(directory))))
    finally:
        lib().lats_topology_destroy(h)


def _same(a, b):
    for name in ("faces", "edges", "bending", "rest_length", "bending_Q", "adj_edges_csr", "adj_bending_csr",
                 "clusters_csr", "colour"):
        u, v = getattr(a, name), getattr(b, name)
        assert u.dtype == v.dtype and u.shape == v.shape, name
        assert np.array_equal(u.view(np.uint8), v.view(np.uint8)), name
    assert (a.num_verts, a.num_edges, a.num_bending, a.num_clusters) == \
           (b.num_verts, b.num_edges, b.num_bending, b.num_clusters)


def test_round_trip_square8k(tmp_path):
    x, f = square8k_mesh()
    built = Topology(x, f)
    d1, d2 = tmp_path / "lib", tmp_path / "py"
    d1.mkdir()
    d2.mkdir()
    _save_with_library(x, f, d1)
    built.save(str(d2))
    for name in FILES:                                   # the two writers agree byte for byte
        assert (d1 / f"{name}.dat").read_bytes() == (d2 / f"{name}.dat").read_bytes(), name
    # the files are the arrays' raw bytes (read_from_binary: count = file size / element size)
    assert (d1 / "vert_adj_edges.dat").stat().st_size == 4 * len(built.adj_edges_csr)
    assert np.array_equal(np.fromfile(d1 / "clusterd_per_vertex_bending.dat", np.uint32), built.clusters_csr)
    _same(Topology.load(str(d1)), built)


def test_square8k_clusters_fixture_id_for_id():
    fixture = np.fromfile(os.path.join(GOLDEN, "square8k_clusters.dat"), np.uint32)
    x, f = square8k_mesh()
    t = Topology(x, f)
    assert np.array_equal(t.clusters_csr, fixture)                           # product builder
    assert np.array_equal(O.Mesh(x, f).csr(O.CSR_CLUSTERS), fixture)         # oracle
    # the fixture itself: packed CSR of 10 clusters (the published schedules)
    # with the published sizes, equal to a brute-force first-fit colouring
    # over spring and bending-pair neighbours, each cluster in ascending ids
    pub = json.load(open(os.path.join(GOLDEN, "square8k_published.json")))
    C = int(fixture[0]) - 1
    assert C == pub["num_clusters"] and fixture[C] == len(fixture)
    clusters = [fixture[fixture[c]:fixture[c + 1]] for c in range(C)]
    assert [len(c) for c in clusters] == [1384, 1226, 1194, 1144, 1087, 971, 757, 430, 85, 4]
    e, b = brute_edges(f.tolist())
    col, _ = first_fit(len(x), e, b)
    got = np.full(len(x), -1)
    for c, verts in enumerate(clusters):
        assert np.all(np.diff(verts.astype(np.int64)) > 0)
        got[verts] = c
    assert got.tolist() == col


def test_loaded_topology_feeds_a_context_description(tmp_path):
    cloth = configs.grid_cloth(16, "corners", seed=4)
    cloth.topology.save(str(tmp_path))
    t = Topology.load(str(tmp_path))
    _same(t, cloth.topology)


@pytest.mark.parametrize("damage", ["missing", "ragged", "csr_header", "csr_offsets", "colour", "cluster_ids"])
def test_damaged_files_are_rejected(tmp_path, damage):
    cloth = configs.grid_cloth(6, "corners", seed=5)
    cloth.topology.save(str(tmp_path))
    path = lambda n: tmp_path / f"{n}.dat"  # noqa: E731
    if damage == "missing":
        path("bending_edges_Q").unlink()
    elif damage == "ragged":
        path("edges").write_bytes(path("edges").read_bytes()[:-4])
    elif damage == "csr_header":
        a = np.fromfile(path("vert_adj_bending_edges"), np.uint32)
        a[0] += 1
        a.tofile(path("vert_adj_bending_edges"))
    elif damage == "csr_offsets":
        a = np.fromfile(path("vert_adj_edges"), np.uint32)
        a[3], a[4] = a[4], a[3]
        a.tofile(path("vert_adj_edges"))
    elif damage == "colour":
        a = np.fromfile(path("per_vertex_bending_cluster_id"), np.uint8)
        a[7] ^= 1
        a.tofile(path("per_vertex_bending_cluster_id"))
    elif damage == "cluster_ids":
        a = np.fromfile(path("clusterd_per_vertex_bending"), np.uint32)
        a[-1] = 10 ** 6
        a.tofile(path("clusterd_per_vertex_bending"))
    with pytest.raises(LatsError) as ei:
        Topology.load(str(tmp_path))
    assert ei.value.code == _lib.LATS_ERR_INVALID_ARGUMENT
