"""Batched cloths: every cloth of the batch == the single-cloth GPU path (and
the oracle) on the same inputs, bit for bit."""
import numpy as np
import pytest

from tests.helpers import oracle_sim
from libatssim_amd import GpuVBD, Params, configs
from libatssim_amd.vbd import GpuBatch

pytestmark = pytest.mark.gpu


def _print_case(name):
    print(f"[case] {name}", flush=True)


def _small_batch(num, seed, nmax=40, iters=20):
    cloths, ks, kb, _ = configs.batch_c2(num_cloths=num, seed=seed, nmin=16, nmax=nmax)
    return cloths, ks, kb, Params(constraint_iter_count=iters)


@pytest.mark.parametrize("threads,lanes", [(64, 1), (256, 1), (1024, 1), (256, 4), (1024, 4), (128, 8), (1024, 8)])
def test_batch_matches_single_cloth(gpu_available, threads, lanes):
    _print_case(f"batch of 6 cloths, {threads} threads per cloth-block, {lanes} lanes per vertex")
    cloths, ks, kb, params = _small_batch(6, seed=11)
    b = GpuBatch(cloths, params, ks, kb)
    b.set_launch(threads, 0)
    b.set_lanes_per_vertex(lanes)
    b.run_frames(3)
    for k, cl in enumerate(cloths):
        p = Params(constraint_iter_count=params.constraint_iter_count, stiffness_stretch_spring=float(ks[k]),
                   stiffness_quadratic_bending=float(kb[k]))
        g = GpuVBD(cl, p)
        for _ in range(3):
            g.physics_step()
        xb, vb = b.state(k)
        xg, vg = g.state()
        assert np.array_equal(xb, xg), k
        assert np.array_equal(vb, vg), k


@pytest.mark.parametrize("staging,threads,lanes", [("none", 256, 4), ("none", 1024, 1), ("none", 128, 8),
                                                    ("some", 256, 4), ("some", 64, 1), ("some", 1024, 8)])
def test_batch_lds_staging_matches_single_cloth(gpu_available, staging, threads, lanes):
    """Positions held in LDS for the frame (cloths up to the staging limit)
    or gathered from HBM give the same bits; "some" stages only the cloths
    of at most 700 vertices, so one launch mixes both paths."""
    _print_case(f"batch of 6 cloths, LDS staging {staging}, {threads} threads, {lanes} lanes per vertex")
    cloths, ks, kb, params = _small_batch(6, seed=13)
    b = GpuBatch(cloths, params, ks, kb)
    limit = 0 if staging == "none" else 700
    b.set_lds_slots(limit)
    assert b.lds_slots()[0] == limit
    if staging == "some":
        sizes = [c.num_verts for c in cloths]
        assert min(sizes) <= limit < max(sizes)
    b.set_launch(threads, 0)
    b.set_lanes_per_vertex(lanes)
    b.run_frames(3)
    for k, cl in enumerate(cloths):
        p = Params(constraint_iter_count=params.constraint_iter_count, stiffness_stretch_spring=float(ks[k]),
                   stiffness_quadratic_bending=float(kb[k]))
        g = GpuVBD(cl, p)
        for _ in range(3):
            g.physics_step()
        xb, vb = b.state(k)
        xg, vg = g.state()
        assert np.array_equal(xb, xg), k
        assert np.array_equal(vb, vg), k


def test_batch_lds_limit(gpu_available):
    _print_case("batch LDS staging limit")
    cloths, ks, kb, params = _small_batch(3, seed=14)
    b = GpuBatch(cloths, params, ks, kb)
    slots, mx = b.lds_slots()
    # default: every cloth that fits one block's LDS (12 bytes per vertex)
    assert slots == min(max(c.num_verts for c in cloths), mx)
    assert 12 * mx <= 160 * 1024
    with pytest.raises(RuntimeError, match="fit in one block's LDS"):
        b.set_lds_slots(mx + 1)
    assert b.lds_slots()[0] == slots


def test_batch_matches_oracle(gpu_available):
    _print_case("batch of 3 cloths vs oracle")
    cloths, ks, kb, params = _small_batch(3, seed=12, nmax=24, iters=15)
    b = GpuBatch(cloths, params, ks, kb)
    b.run_frames(2)
    for k, cl in enumerate(cloths):
        p = Params(constraint_iter_count=15, stiffness_stretch_spring=float(ks[k]),
                   stiffness_quadratic_bending=float(kb[k]))
        o = oracle_sim(cl, p)
        o.step()
        o.step()
        assert np.array_equal(b.state(k)[0], o.state()[0].astype(np.float32)), k


def test_batch_c2_full_size(gpu_available):
    """The 512-cloth workload: autotuned mapping, finite state, spot checks."""
    _print_case("batch_c2 512 cloths: autotune + 2 frames + spot checks")
    cloths, ks, kb, params = configs.batch_c2()
    b = GpuBatch(cloths, params, ks, kb)
    n, nv = b.counts()
    assert n == 512 and nv == sum(c.num_verts for c in cloths)
    x0 = b.state(0)[0]
    t, blocks = b.autotune(frames=1)
    assert t % 64 == 0 and blocks >= 1
    assert np.array_equal(b.state(0)[0], x0)          # autotune restored the state
    b.run_frames(2)
    sizes = [c.num_verts for c in cloths]
    # the largest cloth (beyond one block's LDS: gathered from HBM) and the
    # smallest (staged in LDS) among the spot checks
    assert max(sizes) > b.lds_slots()[1] >= min(sizes)
    for k in (0, 255, 511, int(np.argmax(sizes)), int(np.argmin(sizes))):
        cl = cloths[k]
        p = Params(constraint_iter_count=100, stiffness_stretch_spring=float(ks[k]),
                   stiffness_quadratic_bending=float(kb[k]))
        g = GpuVBD(cl, p)
        g.physics_step()
        g.physics_step()
        assert np.array_equal(b.state(k)[0], g.state()[0]), k
    for k in range(0, 512, 37):
        x, v = b.state(k)
        assert np.all(np.isfinite(x)) and np.all(np.isfinite(v))
