"""Pins the oracle's XPBD path (oracle/vbd_oracle.cpp)
before it checks the GPU solver.  The reference publishes no XPBD run, so:

  * constraint colouring: a set-based numpy restatement of the greedy
    (example2_async_iter.cpp:458-533) on random meshes, plus the defining
    properties (a partition; no two constraints of a cluster share a vertex;
    each constraint sits in the first cluster free of its neighbours);
  * the distance constraint in FP64 against the closed form of one XPBD
    projection of a spring with one pinned end (Macklin et al. 2016, eq. 18:
    dlambda = (-C - alpha~ lambda) / (w + alpha~)), including the rigid
    limit (no compliance: the free end lands at the rest length);
  * the bending constraint: a flat hinge has C = 0 and only falls; a folded
    hinge with stiff bending loses exactly 3/4 of C per iteration (Q is rank
    one, so the projection halves u = sum_i K_i x_i).
"""
import numpy as np
import pytest

from oracle import oracle as O
from libatssim_amd.vbd import make_grid


def brute_colour(elems):
    """First-fit over constraints in id order, one cluster per round."""
    elems = [tuple(int(v) for v in e) for e in elems]
    left = list(range(len(elems)))
    clusters = []
    while left:
        used, cur, rest = set(), [], []
        for k in left:
            if used.isdisjoint(elems[k]):
                cur.append(k)
                used.update(elems[k])
            else:
                rest.append(k)
        clusters.append(cur)
        left = rest
    return clusters


def _sim(x, f, mass=1e-3, fixed=None, iters=1, ks=1e4, kb=1e-2, fp64=False, gravity=(0.0, -9.8, 0.0)):
    m = O.Mesh(np.asarray(x, np.float32), np.asarray(f, np.int32))
    n = len(x)
    fixed = np.zeros(n, np.uint8) if fixed is None else np.asarray(fixed, np.uint8)
    s = O.Sim(m, np.full(n, mass, np.float32), fixed, iter_count=iters, k_stretch=ks, k_bending=kb,
              fp64=fp64, gravity=gravity)
    s._m = m
    return s, m


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_constraint_colouring_matches_brute_force(seed):
    rng = np.random.default_rng(seed)
    x, f = make_grid(7 + seed, 5 + 2 * seed, 0.02, 1e-3, int(rng.integers(1 << 30)))
    s, m = _sim(x, f)
    edges = m.edges().reshape(-1, 2)
    bending = m.bending().reshape(-1, 4)
    for which, elems in ((0, edges), (1, bending)):
        got = [list(c) for c in s.xpbd_clusters(which)]
        assert got == brute_colour(elems)
        seen = sorted(k for c in got for k in c)
        assert seen == list(range(len(elems)))
        for c in got:
            verts = [v for k in c for v in elems[k]]
            assert len(verts) == len(set(verts))


def test_spring_projection_closed_form_fp64():
    L = 0.01
    x = np.array([[0, 0, 0], [L * 1.3, 0, 0], [0, 0, L]], np.float32)   # third vertex only to make a face
    f = np.array([[0, 1, 2]], np.int32)
    # the FP64 oracle sees the FP32 inputs exactly: h = float(1/60), g = float(-9.8), m = float(1e-3)
    h = np.float64(np.float32(1.0 / 60.0))
    g = np.float64(np.float32(-9.8))
    w = 1.0 / np.float64(np.float32(1e-3))
    for ks in (1e3, 1e4, 1e9):
        s, m = _sim(x, f, fixed=[1, 0, 1], iters=1, ks=ks, kb=0.0, fp64=True)
        edges = m.edges().reshape(-1, 2).tolist()
        s.step_xpbd()
        xs, _ = s.state()
        # predict: the free vertex falls by h^2 g; edges (0,1) and (1,2) then project
        # one after the other (different clusters: they share vertex 1)
        p1 = x[1].astype(np.float64) + np.array([0.0, h * (h * g), 0.0])
        alpha = 1.0 / (np.float64(np.float32(ks)) * h * h)
        order = [e for c in s.xpbd_clusters(0) for e in c]
        rest = {k: np.linalg.norm(x[k[0]].astype(np.float64) - x[k[1]]) for k in ((0, 1), (0, 2), (1, 2))}
        for k in order:
            a, b = edges[k]
            if a != 1 and b != 1:
                continue
            other = b if a == 1 else a
            d = p1 - x[other].astype(np.float64)
            l = np.linalg.norm(d)
            C = l - rest[(min(a, b), max(a, b))]
            dl = -C / (w + alpha)
            p1 = p1 + dl * w * d / l
        assert np.allclose(xs[1], p1, rtol=0, atol=1e-12), (ks, xs[1], p1)
        assert np.array_equal(xs[0], x[0].astype(np.float64)) and np.array_equal(xs[2], x[2].astype(np.float64))


def test_rigid_spring_lands_on_rest_length():
    L = 0.02
    x = np.array([[0, 0, 0], [L * 1.5, 0, 0], [0, 0, L]], np.float32)
    f = np.array([[0, 1, 2]], np.int32)
    s, m = _sim(x, f, fixed=[1, 0, 1], iters=1, ks=1e30, kb=0.0, fp64=True, gravity=(0.0, 0.0, 0.0))
    s.step_xpbd()
    xs, _ = s.state()
    # the last projection touching vertex 1 leaves that spring at its rest length
    edges = m.edges().reshape(-1, 2).tolist()
    order = [e for c in s.xpbd_clusters(0) for e in c]
    last = [k for k in order if 1 in edges[k]][-1]
    a, b = edges[last]
    rest = np.linalg.norm(x[a].astype(np.float64) - x[b])
    assert abs(np.linalg.norm(xs[a] - xs[b]) - np.float32(rest)) < 1e-9


def _hinge(fold):
    # two triangles sharing edge (0,1); vertex 3 lifted by `fold`
    return (np.array([[0, 0, 0], [0.02, 0, 0], [0.01, 0.015, 0], [0.01, -0.015, fold]], np.float32),
            np.array([[0, 1, 2], [1, 0, 3]], np.int32))


def test_flat_hinge_does_not_bend():
    """A flat hinge only falls: Q annihilates translations (its K weights sum
    to zero), so C stays 0 and the projection moves nothing (predict applies
    the reference's hard-coded gravity)."""
    x, f = _hinge(0.0)
    s, _ = _sim(x, f, iters=5, ks=0.0, kb=1e2, fp64=True, gravity=(0.0, 0.0, 0.0))
    s.step_xpbd()
    h = np.float64(np.float32(1.0 / 60.0))
    fall = np.array([0.0, h * (h * np.float64(np.float32(-9.8))), 0.0])
    assert np.array_equal(s.state()[0], x.astype(np.float64) + fall)


def test_folded_hinge_bending_closed_form():
    """Q = s K K^T has rank one, so C = s/2 |u|^2 with u = sum_i K_i x_i and
    grad_i C = s K_i u.  One XPBD step without compliance moves each vertex
    by -w K_i u / (2 sum_j w K_j^2), which halves u: C falls by exactly 4 per
    iteration.  The FP32 Q is rank one only to rounding, which leaves a
    small absolute floor (relative 6e-5 after three iterations, growing 4x
    per iteration after that); a wrong sign, a dropped 1/2 or a wrong
    gradient changes the factor itself."""
    x, f = _hinge(0.006)
    m = O.Mesh(x, f)
    _, q = m.rest()
    b = m.bending().reshape(-1, 4)[0]

    def c_bend(pos):
        p = pos[b].astype(np.float64)
        Q = q[0].astype(np.float64)
        return 0.5 * sum(Q[i, j] * p[i] @ p[j] for i in range(4) for j in range(4))

    c0 = c_bend(x)
    assert c0 > 1e-6
    for it in (1, 2, 3):
        s, _ = _sim(x, f, iters=it, ks=0.0, kb=1e9, fp64=True, gravity=(0.0, 0.0, 0.0))
        s.step_xpbd()
        got = c_bend(s.state()[0])
        assert abs(got / (c0 * 0.25 ** it) - 1.0) < 1e-4, (it, got, c0 * 0.25 ** it)
