"""GPU parity of the HIP VBD path against the CPU oracle (FP32, same inputs).

Tolerances: positions 1e-5 relative (max-norm) per frame, energy history
1e-4 relative -- the HIP kernels follow the reference's FP32 operation order
(-ffp-contract=off, IEEE div/sqrt), so the positions are expected to be
bit-identical; the tests check both.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import GOLDEN, oracle_sim, rel_maxnorm, square8k_mesh
from libatssim_amd import GpuVBD, Params, configs, _lib
from libatssim_amd.vbd import Cloth, make_grid

pytestmark = pytest.mark.gpu

NO_ITER = 0xFFFFFFFF
POS_TOL = 1e-5
ENERGY_TOL = 1e-4


def _print_case(name):
    print(f"[case] {name}", flush=True)


def test_c0_frames_bit_exact(gpu_available):
    _print_case("grid_c0 30 frames + energy frame")
    cloth, params, frames, _ = configs.grid_c0()
    g = GpuVBD(cloth, params)
    o = oracle_sim(cloth, params)
    worst = 0.0
    for fr in range(frames):
        g.physics_step()
        o.step()
        xg, vg = g.state()
        xo, vo = o.state()
        worst = max(worst, rel_maxnorm(xg, xo))
        assert rel_maxnorm(xg, xo) <= POS_TOL, fr
        assert np.array_equal(xg, xo.astype(np.float32)), f"frame {fr} not bit-exact"
        assert np.array_equal(vg, vo.astype(np.float32)), f"frame {fr} velocity not bit-exact"
    g.set_energy_bending(False)
    eg = g.physics_step(record_energy=True)
    eo = o.step(record=True)
    assert len(eg) == len(eo) == params.constraint_iter_count + 1
    assert np.abs(eg - eo).max() / np.abs(eo).max() <= ENERGY_TOL
    assert np.array_equal(g.state()[0], o.state()[0].astype(np.float32))


def test_energy_with_bending(gpu_available):
    _print_case("grid_c0 energy with bending")
    cloth, params, _, _ = configs.grid_c0()
    g = GpuVBD(cloth, params)
    o = oracle_sim(cloth, params, energy_bending=True)
    for _ in range(3):
        g.physics_step()
        o.step()
    eg = g.physics_step(record_energy=True)
    eo = o.step(record=True)
    rel = np.abs(eg.astype(np.float64) - eo) / np.abs(eo)
    assert rel.max() <= ENERGY_TOL, rel.max()
    # the solver decreases the incremental potential at every iteration
    assert np.all(np.diff(eg.astype(np.float64)) <= 1e-6 * np.abs(eg).max())


def test_square8k_matches_oracle_and_published(gpu_available):
    _print_case("square8k reference run, 30 frames + energy frame")
    x, f = square8k_mesh()
    cloth = configs.square8k_cloth(x, f)
    params = configs.square8k_params()
    g = GpuVBD(cloth, params)
    g.set_energy_bending(False)          # the reference's CPU energy drops bending
    o = oracle_sim(cloth, params, threads=O.max_threads())
    for fr in range(30):
        g.physics_step()
        o.step()
        if fr % 10 == 9:
            assert rel_maxnorm(g.state()[0], o.state()[0]) <= POS_TOL
    eg = g.physics_step(record_energy=True)
    o.step(record=True)
    assert np.array_equal(g.state()[0], o.state()[0].astype(np.float32))
    pub = np.array(json.load(open(os.path.join(GOLDEN, "square8k_published.json")))["list_sync"])
    assert (np.abs(eg - pub) / pub).max() < 5e-6


def test_run_frames_matches_stepping(gpu_available):
    _print_case("run_frames with per-frame forces == physics_step loop")
    cloth, params, _, _ = configs.grid_c0()
    wind = configs.make_wind(5, cloth.num_verts, 2.0, 99) * 50.0
    a = GpuVBD(cloth, params)
    b = GpuVBD(cloth, params)
    a.run_frames(5, wind)
    for fr in range(5):
        b.set_external_force(wind[fr])
        b.physics_step()
    assert np.array_equal(a.state()[0], b.state()[0])
    assert a.last_elapsed_ms() > 0


def test_c1_long_run_invariants(gpu_available):
    """Full-size run (60 windy frames): finite, pins hold, energy decreases."""
    _print_case("grid_c1 256x256, 60 frames invariants")
    cloth, params, frames, wind = configs.grid_c1(frames=60)
    g = GpuVBD(cloth, params)
    g.run_frames(frames, wind)
    x, v = g.state()
    assert np.all(np.isfinite(x)) and np.all(np.isfinite(v))
    pins = cloth.fixed.astype(bool)
    # the reference's pin is a 1e9 Hessian term with no restoring force
    # (VBD::compute_inertia): the 65 kg cloth hanging on 256 pins drags them
    # by ~1e-3 m over 60 frames, identically in the oracle
    assert np.abs(x[pins] - cloth.x[pins]).max() < 5e-3
    # 100 Gauss-Seidel sweeps do not converge a 256-row chain: the cloth is
    # still falling after 1 s (as in the oracle); the sweeps lower the
    # energy over the frame (single sweeps may not: the clamped Hessian and
    # the pinned-vertex inertia quirk make them non-monotone)
    e = g.physics_step(record_energy=True)
    assert e[-1] < e[0]


def _dispatch_frame(g, ncl, iters, buffer_idx=_lib.LATS_BUFFER_DEFAULT):
    """The reference's single-device frame DAG, issued task by task."""
    g.dispatch(_lib.LATS_FN_ADDITIONAL_ROOT, iter_idx=NO_ITER)
    g.dispatch(_lib.LATS_FN_PREDICT_POSITION, iter_idx=NO_ITER)
    for it in range(iters):
        for c in range(ncl):
            left = _lib.LATS_BUFFER_DEFAULT if (it == 0 and c == 0 and buffer_idx != _lib.LATS_BUFFER_DEFAULT) \
                else _lib.LATS_BUFFER_NONE
            g.dispatch(_lib.LATS_FN_VBD_ALL_IN_ONE, iter_idx=it, cluster_idx=c, buffer_idx=buffer_idx,
                       left_buffer_idx=left)
    # Task(id_xpbd_constraint_last_node, 0, ...) sets cluster_idx = 0 and
    # leaves iter_idx = -1u (launcher.h:633-635)
    g.dispatch(_lib.LATS_FN_CONSTRAINT_LAST_NODE, iter_idx=NO_ITER, buffer_idx=buffer_idx)
    g.dispatch(_lib.LATS_FN_UPDATE_VELOCITY, iter_idx=NO_ITER)
    g.dispatch(_lib.LATS_FN_ADDITIONAL_TERMINAL)
    g.synchronize()


@pytest.mark.parametrize("buffer_idx", [_lib.LATS_BUFFER_DEFAULT, 5])
def test_dispatch_slot_matches_physics_step(gpu_available, buffer_idx):
    _print_case(f"dispatch slot, buffer {buffer_idx:#x}")
    cloth, params, _, _ = configs.grid_c0()
    params = Params(constraint_iter_count=10)
    a = GpuVBD(cloth, params)
    b = GpuVBD(cloth, params)
    ncl = cloth.topology.num_clusters
    b.set_record_energy(True)
    for _ in range(3):
        a.physics_step()
        _dispatch_frame(b, ncl, 10, buffer_idx)
        assert np.array_equal(a.state()[0], b.state()[0])
        assert np.array_equal(a.state()[1], b.state()[1])
    # energy hooks: before iteration 0 of cluster 0, after the last cluster of
    # every iteration, and at the last node
    e = b.energy()
    assert len(e) == 3 * (1 + 10 + 1)
    # rows finalised one at a time survive the history's growth (8 -> 24 -> 56 rows)
    assert np.all(np.isfinite(e)) and np.all(e > 0), e
    for fr in range(3):
        ef = e[12 * fr:12 * fr + 12].astype(np.float64)
        assert np.all(np.diff(ef[:11]) <= 1e-6 * ef.max()), (fr, ef)
        assert ef[11] == ef[10]            # the last node re-evaluates the final iterate


def test_read_and_solve_conflict_merge(gpu_available):
    """Two buffers iterated independently and merged with weight 0.5
    (Core::read_and_solve_conflict) == the numpy restatement of the merge."""
    _print_case("async merge of two buffers")
    cloth, params, _, _ = configs.grid_c0()
    params = Params(constraint_iter_count=1)
    g = GpuVBD(cloth, params)
    g.physics_step()
    g.dispatch(_lib.LATS_FN_PREDICT_POSITION)
    # buffer 1 <- sa_x, one sweep of clusters 0..2 on it
    for c in range(3):
        g.dispatch(_lib.LATS_FN_VBD_ALL_IN_ONE, cluster_idx=c, buffer_idx=1,
                   left_buffer_idx=_lib.LATS_BUFFER_DEFAULT if c == 0 else _lib.LATS_BUFFER_NONE)
    # buffer 2 <- sa_x, clusters 3..6
    for c in range(3, cloth.topology.num_clusters):
        g.dispatch(_lib.LATS_FN_VBD_ALL_IN_ONE, cluster_idx=c, buffer_idx=2,
                   left_buffer_idx=_lib.LATS_BUFFER_DEFAULT if c == 3 else _lib.LATS_BUFFER_NONE)
    # read buffers back through last_node copies
    g.dispatch(_lib.LATS_FN_CONSTRAINT_LAST_NODE, iter_idx=NO_ITER, buffer_idx=1)
    xa = g.state()[0].copy()
    g.dispatch(_lib.LATS_FN_CONSTRAINT_LAST_NODE, iter_idx=NO_ITER, buffer_idx=2)
    xb = g.state()[0].copy()
    # merge 2 <- (input 2, left 1) as the prev hook of a task on buffer 2
    g.dispatch(_lib.LATS_FN_CONSTRAINT_LAST_NODE, iter_idx=NO_ITER, buffer_idx=2, left_buffer_idx=1,
               input_buffer_idxs=(2,),
               is_allocated_to_main_device=True)
    xm = g.state()[0]
    # begin buffer of 2 == predicted positions (copied from sa_x at cluster 3)
    g2 = GpuVBD(cloth, params)
    g2.physics_step()
    g2.dispatch(_lib.LATS_FN_PREDICT_POSITION)
    xp = g2.state()[0]
    sd, ad = xb - xp, xa - xp
    act_a = (sd * sd).sum(1) != 0
    act_b = (ad * ad).sum(1) != 0
    want = xp.copy()
    both = act_a & act_b
    want[both] = xp[both] + np.float32(0.5) * (sd[both] + ad[both])
    want[act_a & ~act_b] = xb[act_a & ~act_b]
    want[~act_a & act_b] = xa[~act_a & act_b]
    assert np.array_equal(xm, want)


def test_det_guard_on_gpu(gpu_available):
    _print_case("det guard: massless isolated vertex")
    base = configs.grid_cloth(4, "corners", seed=2)
    x = np.vstack([base.x, [[0.5, 0.5, 0.0]]]).astype(np.float32)
    cloth = Cloth(x, base.faces, np.append(base.mass, 0.0), np.append(base.fixed, 0))
    params = Params(constraint_iter_count=10)
    g = GpuVBD(cloth, params)
    o = oracle_sim(cloth, params)
    for _ in range(3):
        g.physics_step()
        o.step()
    xg = g.state()[0]
    assert np.all(np.isfinite(xg))
    assert np.array_equal(xg, o.state()[0].astype(np.float32))


@pytest.mark.parametrize("n", [2, 3, 5])
def test_tiny_cloths(gpu_available, n):
    _print_case(f"tiny {n}x{n} cloth")
    cloth = configs.grid_cloth(n, "top", seed=n)
    params = Params(constraint_iter_count=7)
    g = GpuVBD(cloth, params)
    o = oracle_sim(cloth, params)
    for _ in range(4):
        g.physics_step()
        o.step()
    assert np.array_equal(g.state()[0], o.state()[0].astype(np.float32))


def test_substeps(gpu_available):
    _print_case("num_substep = 3")
    cloth, _, _, _ = configs.grid_c0()
    params = Params(num_substep=3, constraint_iter_count=5)
    g = GpuVBD(cloth, params)
    o = oracle_sim(cloth, params)
    for _ in range(3):
        g.physics_step()
        o.step()
    assert np.array_equal(g.state()[0], o.state()[0].astype(np.float32))


def test_host_signal_and_markers(gpu_available):
    _print_case("host signal / markers")
    cloth, params, _, _ = configs.grid_c0()
    g = GpuVBD(cloth, Params(constraint_iter_count=3))
    L = _lib.lib()
    _lib.check(L.lats_vbd_record_marker(g.handle, 1))
    _lib.check(L.lats_vbd_host_signal(g.handle, 7))
    _lib.check(L.lats_vbd_stream_wait_host(g.handle, 7))
    g.dispatch(_lib.LATS_FN_PREDICT_POSITION)
    _lib.check(L.lats_vbd_record_marker(g.handle, 2))
    _lib.check(L.lats_vbd_wait_marker(g.handle, 2))
    import ctypes
    ms = ctypes.c_float()
    _lib.check(L.lats_vbd_marker_elapsed_ms(g.handle, 1, 2, ctypes.byref(ms)))
    assert ms.value >= 0.0


@pytest.mark.parametrize("blocks", [0, 7, 200])
def test_persistent_strategy_bit_exact(gpu_available, blocks):
    """One cooperative launch per run (grid barrier per colour) == graph path."""
    _print_case(f"persistent strategy, {blocks} blocks, grid_c0 + wind")
    cloth, params, _, _ = configs.grid_c0()
    wind = configs.make_wind(6, cloth.num_verts, 2.0, 5) * 30.0
    a = GpuVBD(cloth, params)
    a.set_strategy(_lib.LATS_STRATEGY_GRAPH)
    b = GpuVBD(cloth, params)
    b.set_strategy(_lib.LATS_STRATEGY_PERSISTENT, persistent_blocks=blocks)
    a.run_frames(6, wind)
    b.run_frames(6, wind)
    assert np.array_equal(a.state()[0], b.state()[0])
    assert np.array_equal(a.state()[1], b.state()[1])
    # physics_step after a persistent run continues the same frame sequence
    a.set_external_force(wind[0])
    b.set_external_force(wind[0])
    a.physics_step()
    b.physics_step()
    assert np.array_equal(a.state()[0], b.state()[0])


def test_persistent_c1_and_autotune(gpu_available):
    _print_case("persistent strategy on grid_c1 + autotune")
    cloth, params, _, wind = configs.grid_c1(frames=3)
    a = GpuVBD(cloth, params)
    a.set_strategy(_lib.LATS_STRATEGY_GRAPH)
    b = GpuVBD(cloth, params)
    x_before = b.state()[0]
    st, nb = b.autotune(frames=1)
    assert st in (_lib.LATS_STRATEGY_GRAPH, _lib.LATS_STRATEGY_PERSISTENT, _lib.LATS_STRATEGY_FLOW,
                  _lib.LATS_STRATEGY_PATCH)
    assert np.array_equal(b.state()[0], x_before)          # autotune restores the state
    # the reported candidates: the kept configuration is the fastest one
    res = b.autotune_results()
    assert res.shape[1] == 6 and len(res) >= 8 and (res[:, 4] > 0).all()
    win = res[np.argmin(res[:, 4])]
    assert int(win[0]) == st and int(win[2]) == b.lanes_per_vertex()
    if st == _lib.LATS_STRATEGY_GRAPH:
        assert int(win[3]) == b.regions() and int(win[5]) == b.colour_block()
    if st == _lib.LATS_STRATEGY_PATCH:
        info = b.patch_info()
        assert int(win[1]) == info["patches"] == nb and int(win[5]) == info["threads"]
    # patch candidates are timed on this cloth (65,536 vertices)
    assert (res[:, 0] == _lib.LATS_STRATEGY_PATCH).sum() >= 1
    b.set_strategy(_lib.LATS_STRATEGY_PERSISTENT)
    a.run_frames(3, wind)
    b.run_frames(3, wind)
    assert np.array_equal(a.state()[0], b.state()[0])


@pytest.mark.parametrize("lanes", [1, 4, 8, 16])
def test_region_mapping_bit_exact(gpu_available, lanes):
    """Colour passes with blocks mapped to 8 cloth bands == the linear mapping."""
    _print_case(f"region block mapping, {lanes} lanes, grid_c1 + wind, 3 frames")
    cloth, params, _, wind = configs.grid_c1(frames=3)
    a = GpuVBD(cloth, params)
    a.set_strategy(_lib.LATS_STRATEGY_GRAPH)
    a.set_lanes_per_vertex(lanes)
    b = GpuVBD(cloth, params)
    b.set_strategy(_lib.LATS_STRATEGY_GRAPH)
    b.set_lanes_per_vertex(lanes)
    b.set_regions(8)
    assert b.regions() == 8
    a.run_frames(3, wind)
    b.run_frames(3, wind)
    assert np.array_equal(a.state()[0], b.state()[0])
    assert np.array_equal(a.state()[1], b.state()[1])


@pytest.mark.parametrize("n,regions", [(2, 8), (3, 8), (5, 3), (17, 7)])
def test_region_mapping_small_cloths(gpu_available, n, regions):
    """Bands with no vertex of a colour (cloths smaller than the band count)."""
    _print_case(f"region block mapping, {n}x{n} cloth, {regions} regions")
    cloth = configs.grid_cloth(n, "top", seed=n)
    params = Params(constraint_iter_count=10)
    b = GpuVBD(cloth, params)
    b.set_strategy(_lib.LATS_STRATEGY_GRAPH)
    b.set_lanes_per_vertex(8)
    b.set_regions(regions)
    o = oracle_sim(cloth, params)
    for _ in range(3):
        b.physics_step()
        o.step()
    assert np.array_equal(b.state()[0], o.state()[0].astype(np.float32))
    with pytest.raises(_lib.LatsError):
        b.set_regions(0)
    with pytest.raises(_lib.LatsError):
        b.set_regions(9)
    assert b.regions() == regions


@pytest.mark.parametrize("tiles,lanes", [(1, 8), (4, 1), (4, 8), (16, 4), (32, 8), (32, 16)])
def test_flow_strategy_bit_exact(gpu_available, tiles, lanes):
    """Tiles handing colour passes to their neighbours == graph path, bit for bit."""
    _print_case(f"flow strategy, {tiles} tiles, {lanes} lanes, grid_c0 + wind")
    cloth, params, _, _ = configs.grid_c0()
    wind = configs.make_wind(6, cloth.num_verts, 2.0, 5) * 30.0
    a = GpuVBD(cloth, params)
    a.set_strategy(_lib.LATS_STRATEGY_GRAPH)
    b = GpuVBD(cloth, params)
    b.set_strategy(_lib.LATS_STRATEGY_FLOW, persistent_blocks=tiles)
    b.set_lanes_per_vertex(lanes)
    a.run_frames(6, wind)
    b.run_frames(6, wind)
    assert np.array_equal(a.state()[0], b.state()[0])
    assert np.array_equal(a.state()[1], b.state()[1])
    # single frames, a graph frame in between (counters carry over), then flow again
    for k, strategy in enumerate((_lib.LATS_STRATEGY_FLOW, _lib.LATS_STRATEGY_GRAPH, _lib.LATS_STRATEGY_FLOW)):
        b.set_strategy(strategy)
        a.set_external_force(wind[k])
        b.set_external_force(wind[k])
        a.physics_step()
        b.physics_step()
        assert np.array_equal(a.state()[0], b.state()[0]), k
        assert np.array_equal(a.state()[1], b.state()[1]), k


def test_flow_c1_bit_exact(gpu_available):
    _print_case("flow strategy on grid_c1 (default tiles), 4 windy frames")
    cloth, params, _, wind = configs.grid_c1(frames=4)
    a = GpuVBD(cloth, params)
    a.set_strategy(_lib.LATS_STRATEGY_GRAPH)
    a.set_lanes_per_vertex(8)
    b = GpuVBD(cloth, params)
    b.set_strategy(_lib.LATS_STRATEGY_FLOW)
    b.set_lanes_per_vertex(8)
    a.run_frames(4, wind)
    b.run_frames(4, wind)
    assert np.array_equal(a.state()[0], b.state()[0])
    assert np.array_equal(a.state()[1], b.state()[1])


def test_flow_rejects_obstacles_and_oversized_grids(gpu_available):
    _print_case("flow strategy argument checks")
    cloth, params, _, _ = configs.grid_c0()
    b = GpuVBD(cloth, params)
    b.set_strategy(_lib.LATS_STRATEGY_FLOW, persistent_blocks=100000)
    with pytest.raises(_lib.LatsError):
        b.run_frames(1)


def test_context_from_saved_topology_bit_exact(gpu_available, tmp_path):
    """square8K: topology saved to disk and loaded back (lats_topology_save /
    _load) drives a context whose frames equal the built topology's bit for
    bit, and the oracle's."""
    _print_case("square8k from a saved topology, 5 frames")
    from libatssim_amd import Topology
    x, f = square8k_mesh()
    cloth = configs.square8k_cloth(x, f)
    params = configs.square8k_params()
    cloth.topology.save(str(tmp_path))
    loaded = Cloth(cloth.x, cloth.faces, cloth.mass, cloth.fixed, topology=Topology.load(str(tmp_path)))
    a = GpuVBD(cloth, params)
    b = GpuVBD(loaded, params)
    o = oracle_sim(cloth, params, threads=O.max_threads())
    for _ in range(5):
        a.physics_step()
        b.physics_step()
        o.step()
    xa, va = a.state()
    xb, vb = b.state()
    assert np.array_equal(xa, xb) and np.array_equal(va, vb)
    assert np.array_equal(xb, o.state()[0].astype(np.float32))
