"""The C-ABI library loads on a host without a GPU and exports every symbol
declared in include/libatssim_hip.h; calls that need no device work."""
import ctypes
import os
import re

import numpy as np

from libatssim_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "libatssim_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lats_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert set(declared_symbols()) == set(_lib.SIGNATURES)


def test_errors_are_reported():
    L = _lib.lib()
    rc = L.lats_vbd_create(None, None, None)
    assert rc == _lib.LATS_ERR_INVALID_ARGUMENT
    assert b"null" in L.lats_last_error()


def test_create_without_device_fails_loudly():
    """No silent CPU path: creating a context on a host without a GPU fails."""
    import pytest
    hip = ctypes.CDLL("libamdhip64.so")
    n = ctypes.c_int(0)
    if hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0:
        pytest.skip("a GPU is present")
    from libatssim_amd import configs, GpuVBD
    cloth, params, _, _ = configs.grid_c0()
    with pytest.raises(_lib.LatsError) as ei:
        GpuVBD(cloth, params)
    assert ei.value.code == _lib.LATS_ERR_NO_DEVICE


def _with_clusters(cloth, clusters):
    """The cloth with its colour clusters replaced by `clusters` (lists of ids)."""
    import copy
    from libatssim_amd.vbd import Cloth
    t = copy.copy(cloth.topology)
    n = len(clusters)
    offs = [n + 1]
    for c in clusters:
        offs.append(offs[-1] + len(c))
    t.clusters_csr = np.concatenate([np.array(offs, np.uint32)] + [np.asarray(c, np.uint32) for c in clusters])
    t.num_clusters = n
    return Cloth(cloth.x, cloth.faces, cloth.mass, cloth.fixed, topology=t)


def test_dependent_clusters_are_rejected():
    """A colour pass updates its cluster concurrently: a colouring whose
    cluster holds both ends of a spring, or two vertices of one bending pair,
    is refused before any device work (also on a host without a GPU)."""
    import pytest
    from libatssim_amd import GpuVBD, Params, configs
    cloth = configs.grid_cloth(6, "corners", seed=3)
    good = [list(c) for c in cloth.topology.clusters()]
    p = Params(constraint_iter_count=2)
    # one cluster with every vertex
    with pytest.raises(_lib.LatsError) as ei:
        GpuVBD(_with_clusters(cloth, [list(range(cloth.num_verts))]), p)
    assert ei.value.code == _lib.LATS_ERR_INVALID_ARGUMENT and "not independent" in str(ei.value)
    # move one vertex into the cluster of a spring neighbour
    a, b = (int(v) for v in cloth.topology.edges[0])
    bad = [[v for v in c if v != a] for c in good]
    ca = next(k for k, c in enumerate(good) if b in c)
    bad[ca] = sorted(bad[ca] + [a])
    bad = [c for c in bad if c]
    with pytest.raises(_lib.LatsError) as ei:
        GpuVBD(_with_clusters(cloth, bad), p)
    assert ei.value.code == _lib.LATS_ERR_INVALID_ARGUMENT and "spring" in str(ei.value)
    # the two opposite vertices of a bending pair (not joined by a spring) in
    # one cluster: move o0 into o1's cluster where no spring neighbour of o0 is
    colour = np.zeros(cloth.num_verts, np.int64)
    for k, c in enumerate(good):
        colour[np.asarray(c, np.int64)] = k
    nbr = {v: set() for v in range(cloth.num_verts)}
    for a, b in cloth.topology.edges:
        nbr[int(a)].add(int(b))
        nbr[int(b)].add(int(a))
    o0, c1 = next((int(q[u]), int(colour[q[w]])) for q in cloth.topology.bending for u, w in ((2, 3), (3, 2))
                  if all(colour[n] != colour[q[w]] for n in nbr[int(q[u])]))
    bad = [[v for v in c if v != o0] for c in good]
    bad[c1] = sorted(bad[c1] + [o0])
    bad = [c for c in bad if c]
    with pytest.raises(_lib.LatsError) as ei:
        GpuVBD(_with_clusters(cloth, bad), p)
    assert ei.value.code == _lib.LATS_ERR_INVALID_ARGUMENT and "bending pair" in str(ei.value)
