"""The patch strategy (several colour passes per launch, halo updates
recomputed in LDS) against the graph strategy and the oracle.

The patch launches read one position buffer and write the other, and every
update is the graph strategy's arithmetic on the same inputs, so positions and
velocities are expected to be bit-identical for every geometry (patches,
colours per launch, threads; patch updates always run on 8 lanes, the graph
strategy compared against runs on 8, 4 or 16).  The layout itself -- bounds of every
index a launch computes, and Gauss-Seidel order under emulation -- is checked
on the CPU by tests/cpp/test_patch_layout.hip.
"""
import numpy as np
import pytest

from tests.helpers import oracle_sim, square8k_mesh
from libatssim_amd import GpuVBD, Params, configs, _lib
from libatssim_amd.vbd import Cloth, make_grid, make_icosphere

pytestmark = pytest.mark.gpu


def _print_case(name):
    print(f"[case] {name}", flush=True)


def _pair(cloth, params, patches=0, colours=0, threads=1024, lanes=8):
    """A graph-strategy context with `lanes` lanes per vertex and a patch one."""
    a = GpuVBD(cloth, params)
    a.set_strategy(_lib.LATS_STRATEGY_GRAPH)
    a.set_lanes_per_vertex(lanes)
    b = GpuVBD(cloth, params)
    b.set_lanes_per_vertex(lanes)   # (patch updates run on 8 lanes regardless)
    b.set_patches(patches, colours, threads)
    b.set_strategy(_lib.LATS_STRATEGY_PATCH)
    assert b.patch_info()["lanes"] == 8
    return a, b


def test_patch_c0_matches_oracle(gpu_available):
    _print_case("patch strategy, 32x32 c0 cloth, 10 frames vs oracle")
    cloth, params, _, _ = configs.grid_c0()
    g = GpuVBD(cloth, params)
    g.set_patches(16, 0, 256)
    g.set_strategy(_lib.LATS_STRATEGY_PATCH)
    o = oracle_sim(cloth, params)
    for fr in range(10):
        g.physics_step()
        o.step()
        xg, vg = g.state()
        xo, vo = o.state()
        assert np.array_equal(xg, xo.astype(np.float32)), f"frame {fr}"
        assert np.array_equal(vg, vo.astype(np.float32)), f"frame {fr}"
    info = g.patch_info()
    assert info["patches"] == 16 and info["launches_per_sweep"] == 1
    assert info["updates_per_sweep"] >= cloth.num_verts


@pytest.mark.parametrize("patches,colours,threads,lanes", [
    (0, 0, 1024, 8), (128, 0, 1024, 8), (256, 4, 512, 4), (512, 2, 256, 16), (256, 1, 1024, 8), (64, 3, 1024, 16),
])
def test_patch_c1_matches_graph(gpu_available, patches, colours, threads, lanes):
    _print_case(f"patch strategy P={patches} F={colours} B={threads} G={lanes}, grid_c1 + wind, 4 frames")
    cloth, params, _, wind = configs.grid_c1(frames=4)
    a, b = _pair(cloth, params, patches, colours, threads, lanes)
    a.run_frames(4, wind)
    b.run_frames(4, wind)
    assert np.array_equal(a.state()[0], b.state()[0])
    assert np.array_equal(a.state()[1], b.state()[1])


def test_patch_square8k_and_energy(gpu_available):
    """The reference's own mesh (irregular, 10 colours): positions and the
    energy history of a recorded frame equal the graph strategy's."""
    _print_case("patch strategy, square8K, 3 frames + recorded energy")
    x, f = square8k_mesh()
    cloth = configs.square8k_cloth(x, f)
    params = configs.square8k_params()
    for colours in (0, 3):
        a, b = _pair(cloth, params, 64, colours, 512, 8)
        for _ in range(2):
            a.physics_step()
            b.physics_step()
        assert np.array_equal(a.state()[0], b.state()[0])
        ea = a.physics_step(record_energy=True)
        eb = b.physics_step(record_energy=True)
        assert np.array_equal(ea, eb)
        assert np.array_equal(a.state()[0], b.state()[0])


def test_patch_iterate_and_odd_grid(gpu_available):
    """Frame pieces (predict, k sweeps, update) with an odd launch count per
    sweep, on a 45x37 cloth whose last patches are ragged."""
    _print_case("patch strategy, frame pieces, 45x37 cloth")
    xg, fg = make_grid(45, 37, seed=3)
    cloth = Cloth(xg, fg, np.float32(1e-3), np.zeros(45 * 37, np.uint8))
    params = Params(constraint_iter_count=30)
    a, b = _pair(cloth, params, 24, 3, 256, 8)   # 3 launches per sweep
    for g in (a, b):
        for sweeps in (5, 7, 18):
            g.frame_begin()
            g.iterate(sweeps)
            g.frame_end()
    assert np.array_equal(a.state()[0], b.state()[0])
    assert np.array_equal(a.state()[1], b.state()[1])


def test_patch_contacts_match_graph(gpu_available):
    """Obstacle contacts (detected each frame, solved inside the patch
    launches) on a 96x96 cloth dropped onto the sphere."""
    _print_case("patch strategy, 96x96 cloth on the sphere, 6 frames")
    n, spacing, radius = 96, 0.002, 0.3
    x, f = configs.horizontal_grid(n, spacing, radius + 0.015, seed=91)
    cloth = Cloth(x, f, np.float32(1e-3), np.zeros(n * n, np.uint8))
    ox, of = make_icosphere(5, radius)
    params = Params(constraint_iter_count=20)
    a, b = _pair(cloth, params, 64, 0, 512, 8)
    for g in (a, b):
        g.set_obstacle(ox, of, 0.002, 1e5, 0.025, 0.004)
    for _ in range(6):
        a.physics_step()
        b.physics_step()
    ca, _ = a.contacts()
    assert ca.sum() > 0
    assert np.array_equal(a.state()[0], b.state()[0])
    assert np.array_equal(a.state()[1], b.state()[1])


def test_patch_rejects_bad_geometry(gpu_available):
    cloth, params, _, _ = configs.grid_c0()
    g = GpuVBD(cloth, params)
    with pytest.raises(_lib.LatsError):
        g.set_patches(12, 0, 1024)          # not a multiple of 8
    with pytest.raises(_lib.LatsError):
        g.set_patches(8, 0, 384)            # unsupported block size
    g.set_patches(8, 0, 1024)
    g.set_strategy(_lib.LATS_STRATEGY_PATCH)
    g.physics_step()
    assert g.patch_info()["patches"] == 8
