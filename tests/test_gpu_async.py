"""The asynchronous frame on the HIP slot.

The launch sequence of the async schedule (HEFT + async post-pass, run in
LaunchModeSequeceHetero order by tests/cpp/bin/test_async_schedule) is
replayed through lats_vbd_dispatch on one GPU context -- every task, whichever
processor it was scheduled on, with its buffers, merges and copies -- and
compared with the oracle's replay of the same sequence: positions bit for
bit, energies within 1e-4 (the GPU sums them in FP32), and against the
reference's published async history (documents/example2_async_convergence.py:82-85).
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import GOLDEN, oracle_sim, square8k_mesh
from tests.test_async import launch_sequence, replay_on_oracle
from libatssim_amd import GpuVBD, configs

pytestmark = pytest.mark.gpu

ENERGY_TOL = 1e-4


def _print_case(name):
    print(f"[case] {name}", flush=True)


def replay_on_gpu(g, seq):
    for proc, fn, it, cl, buf, left, right, main, ins in seq:
        g.dispatch(fn, iter_idx=it, cluster_idx=cl, buffer_idx=buf, left_buffer_idx=left, right_buffer_idx=right,
                   input_buffer_idxs=ins, is_allocated_to_main_device=main)
    g.synchronize()


def test_square8k_async_frame_on_gpu(gpu_available):
    _print_case("square8k: 30 frames, then the async frame replayed on the GPU slot")
    x, f = square8k_mesh()
    cloth = configs.square8k_cloth(x, f)
    params = configs.square8k_params()
    seq = launch_sequence(100, 10)
    g = GpuVBD(cloth, params)
    g.set_energy_bending(False)          # the reference's CPU energy drops bending
    o = oracle_sim(cloth, params, threads=O.max_threads())
    for _ in range(30):
        g.physics_step()
        o.step()
    g.set_record_energy(True)
    replay_on_gpu(g, seq)
    eg = g.energy()
    eo = replay_on_oracle(o, seq, record=True)
    assert np.array_equal(g.state()[0], o.state()[0].astype(np.float32))
    assert len(eg) == len(eo) == 102
    assert (np.abs(eg.astype(np.float64) - eo) / np.abs(eo)).max() <= ENERGY_TOL
    pub = np.array(json.load(open(os.path.join(GOLDEN, "square8k_published.json")))["list_async"])
    assert (np.abs(eg - pub) / pub).max() < 5e-6


@pytest.mark.parametrize("iters", [10, 37])
def test_grid_async_frames_on_gpu(gpu_available, iters):
    """Three async frames on a windless 32x32 grid: GPU replay == oracle replay."""
    _print_case(f"grid_c0, async schedule with {iters} iterations, 3 frames")
    cloth, _, _, _ = configs.grid_c0()
    from libatssim_amd import Params
    params = Params(constraint_iter_count=iters)
    ncl = cloth.topology.num_clusters
    seq = launch_sequence(iters, ncl)
    g = GpuVBD(cloth, params)
    o = oracle_sim(cloth, params)
    for _ in range(3):
        replay_on_gpu(g, seq)
        replay_on_oracle(o, seq, record=False)
        assert np.array_equal(g.state()[0], o.state()[0].astype(np.float32))
        assert np.array_equal(g.state()[1], o.state()[1].astype(np.float32))


HETERO_BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "bin", "test_hetero_vbd")


@pytest.mark.parametrize("n,iters,frames", [(24, 20, 2), (48, 100, 2)])
def test_hetero_launch_cpu_oracle_plus_gpu(gpu_available, n, iters, frames):
    """LaunchModeHetero with the CPU oracle as processor 0 and the HIP context
    as processor 1 (concurrent, host-value waits and markers),
    LaunchModeProgressiveHetero (GPU events queued three at a time) and
    LaunchModeFakeHetero, each == the two-memory oracle replay bit for bit
    (tests/cpp/test_hetero_vbd.cpp)."""
    import subprocess
    _print_case(f"LaunchModeHetero CPU oracle + GPU, {n}x{n}, {iters} iterations, {frames} frames")
    if not os.path.exists(HETERO_BIN):
        pytest.fail(f"{HETERO_BIN} not built (make -C libatssim_amd/csrc)")
    r = subprocess.run([HETERO_BIN, str(n), str(iters), str(frames)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-3000:])
    assert r.returncode == 0 and "PASS" in r.stdout, r.stdout + r.stderr
    assert r.stdout.count("hetero == two-memory replay: bit-exact, fake-hetero: bit-exact, progressive: bit-exact") \
        == frames
