"""CPU model of the flow strategy's hand-off rule (k_vbd_flow): a tile
starts phase q once every neighbouring tile completed q - 1.  Under random
interleavings of phase starts and ends, every vertex update must read each
neighbour at exactly the version the colour-by-colour order gives, no read
may overlap the write of the same vertex, and the run must not deadlock.
The tile ranges and neighbour sets are built as run_flow builds them."""
import random

import numpy as np
import pytest

from libatssim_amd import configs


def _layout(n, tile_verts):
    cloth = configs.grid_cloth(n, "top", 7)
    t = cloth.topology
    V, C = t.num_verts, t.num_clusters
    colour = np.zeros(V, int)
    for c, vs in enumerate(t.clusters()):
        colour[vs] = c
    tiles = (V + tile_verts - 1) // tile_verts
    tile = np.arange(V) // tile_verts
    reads = [set() for _ in range(V)]
    nb = [set() for _ in range(tiles)]
    for a, b in t.edges:
        reads[a].add(b)
        reads[b].add(a)
    for q in t.bending:
        for i in q:
            for j in q:
                if i != j:
                    reads[i].add(j)
    for group in list(t.edges) + list(t.bending):
        for i in group:
            for j in group:
                if tile[i] != tile[j]:
                    nb[tile[i]].add(tile[j])
    # per (tile, colour) ranges over the colour's ascending vertex ids
    cover = np.zeros(V, int)
    for c in range(C):
        ids = np.sort(np.where(colour == c)[0])
        for tt in range(tiles):
            lo = np.searchsorted(ids, tt * tile_verts)
            hi = np.searchsorted(ids, (tt + 1) * tile_verts)
            assert np.all(tile[ids[lo:hi]] == tt)
            cover[ids[lo:hi]] += 1
    assert np.all(cover == 1)
    return V, C, tiles, tile, colour, reads, nb


def _run(V, C, tiles, tile, colour, reads, nb, iters, rng):
    nph = iters * C
    progress = [0] * tiles
    running = [None] * tiles
    version = np.zeros(V, int)
    writing = set()
    steps = 0
    while min(progress) < nph:
        steps += 1
        assert steps < 200000, "no progress: deadlock"
        t = rng.randrange(tiles)
        if running[t] is None:
            q = progress[t] + 1
            if q > nph or any(progress[u] < q - 1 for u in nb[t]):
                continue
            running[t] = q
            c, k = (q - 1) % C, (q - 1) // C
            for v in np.where((tile == t) & (colour == c))[0]:
                for u in reads[v]:
                    assert u not in writing, "read overlaps a write"
                    assert version[u] == (k + 1 if colour[u] < c else k)
                writing.add(v)
        else:
            c = (running[t] - 1) % C
            for v in np.where((tile == t) & (colour == c))[0]:
                version[v] += 1
                writing.discard(v)
            progress[t] = running[t]
            running[t] = None
    assert np.all(version == iters)


@pytest.mark.parametrize("n,tile_verts", [(12, 12), (10, 7), (16, 20)])
def test_flow_handoff_rule(n, tile_verts):
    layout = _layout(n, tile_verts)
    for trial in range(10):
        _run(*layout, iters=3, rng=random.Random(trial))
