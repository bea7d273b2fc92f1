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
    for k in (0, 255, 511):
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
