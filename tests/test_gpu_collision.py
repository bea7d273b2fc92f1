"""Obstacle contact on the GPU (spatial-hash detection + penalty term in the
VBD solve) == the oracle's brute-force detection and contact term."""
import numpy as np
import pytest

from tests.helpers import oracle_sim, rel_maxnorm
from libatssim_amd import GpuVBD, Params, configs, _lib
from libatssim_amd.vbd import Cloth, make_icosphere

pytestmark = pytest.mark.gpu


def _print_case(name):
    print(f"[case] {name}", flush=True)


def _small_scene(n=24, spacing=0.01, drop=0.015, radius=0.1, level=3):
    x, f = configs.horizontal_grid(n, spacing, radius + drop, seed=77)
    cloth = Cloth(x, f, np.float32(1e-3), np.zeros(n * n, np.uint8))
    ox, of = make_icosphere(level, radius)
    return cloth, (ox, of), dict(thickness=0.002, stiffness=1e5, radius=0.02, cell=0.04)


# 23 x 23 and 25 x 25: vertex counts that are not a multiple of the 32
# vertices a detection block covers, so its last block is partial
@pytest.mark.parametrize("lanes,strategy,n", [(1, "graph", 24), (8, "graph", 24), (8, "graph", 23), (8, "graph", 25),
                                              (4, "persistent", 24)])
def test_contacts_match_oracle(gpu_available, lanes, strategy, n):
    _print_case(f"cloth {n}x{n} on sphere vs oracle, lanes {lanes}, {strategy}")
    cloth, (ox, of), ct = _small_scene(n=n)
    params = Params(constraint_iter_count=20)
    g = GpuVBD(cloth, params)
    g.set_lanes_per_vertex(lanes)
    g.set_strategy(_lib.LATS_STRATEGY_GRAPH if strategy == "graph" else _lib.LATS_STRATEGY_PERSISTENT)
    g.set_obstacle(ox, of, ct["thickness"], ct["stiffness"], ct["radius"], ct["cell"])
    o = oracle_sim(cloth, params, threads=4)
    o.set_obstacle(ox, of, ct["thickness"], ct["stiffness"], ct["radius"])
    touched = 0
    for fr in range(8):
        g.physics_step()
        o.step()
        cg, fg = g.contacts()
        co, fo = o.contacts()
        assert np.array_equal(cg, co), fr
        assert np.array_equal(fg, fo), fr
        touched += int((cg > 0).sum())
        xg, xo = g.state()[0], o.state()[0]
        assert rel_maxnorm(xg, xo) <= 1e-5, fr
        assert np.array_equal(xg, xo.astype(np.float32)), fr
    assert touched > 0


def test_contact_energy_matches_oracle(gpu_available):
    _print_case("contact energy history")
    cloth, (ox, of), ct = _small_scene()
    params = Params(constraint_iter_count=20)
    g = GpuVBD(cloth, params)
    g.set_obstacle(ox, of, ct["thickness"], ct["stiffness"], ct["radius"], ct["cell"])
    o = oracle_sim(cloth, params, threads=4, energy_bending=True)
    o.set_obstacle(ox, of, ct["thickness"], ct["stiffness"], ct["radius"])
    for _ in range(4):
        g.physics_step()
        o.step()
    eg = g.physics_step(record_energy=True)
    eo = o.step(record=True)
    assert np.abs(eg.astype(np.float64) - eo).max() / np.abs(eo).max() <= 1e-4


@pytest.mark.parametrize("lanes,threads", [(4, 256), (8, 256), (16, 256), (1, 64), (8, 64), (4, 128), (16, 128)])
def test_lanes_per_vertex_bit_exact(gpu_available, lanes, threads):
    """G lanes per vertex (terms summed in incidence order) in blocks of any
    size == one thread per vertex in 256-thread blocks."""
    _print_case(f"lanes per vertex {lanes}, {threads}-thread blocks: grid_c0 and grid_c1")
    cloth, params, _, _ = configs.grid_c0()
    a = GpuVBD(cloth, params)
    b = GpuVBD(cloth, params)
    b.set_lanes_per_vertex(lanes)
    b.set_colour_block(threads)
    for _ in range(3):
        a.physics_step()
        b.physics_step()
    assert np.array_equal(a.state()[0], b.state()[0])
    cloth, params, _, wind = configs.grid_c1(frames=2)
    a = GpuVBD(cloth, params)
    b = GpuVBD(cloth, params)
    b.set_lanes_per_vertex(lanes)
    b.set_colour_block(threads)
    b.set_regions(8)
    a.run_frames(2, wind)
    b.run_frames(2, wind)
    assert np.array_equal(a.state()[0], b.state()[0])
    e = b.physics_step(record_energy=True)
    assert len(e) == 101


def test_sphere_c4_full_size(gpu_available):
    """1024x1024 cloth dropped on the sphere: 20 frames, contacts appear, no blow-up."""
    _print_case("sphere_c4 1024x1024, 20 frames")
    cloth, params, (ox, of), ct = configs.sphere_scene()
    g = GpuVBD(cloth, params)
    g.set_obstacle(ox, of, ct["thickness"], ct["stiffness"], ct["radius"], ct["cell"])
    g.set_lanes_per_vertex(8)
    g.run_frames(20)
    x, v = g.state()
    assert np.all(np.isfinite(x)) and np.all(np.isfinite(v))
    counts, _ = g.contacts()
    assert counts.sum() > 0
    # vertices in contact sit on or above the sphere surface minus the thickness band
    r = np.linalg.norm(x[counts > 0].astype(np.float64), axis=1)
    print("contact vertices", int((counts > 0).sum()), "min radius", r.min())
    assert r.min() > 0.3 - 0.01


@pytest.mark.parametrize("cell", [0.05, 0.004])
def test_contacts_match_oracle_dense_cells(gpu_available, cell):
    """The sphere_c4 sphere and contact parameters (level-5 icosphere, 2.5 cm
    query radius) on a 96x96 piece of the cloth dropped 1.5 cm, with 5 cm
    cells (~100 faces per cell, so all 8 lanes of the detection kernel find
    candidates and their nearest lists are merged) and with sphere_c4's own
    4 mm cells (2x the edge length): contact sets and positions equal the
    oracle's brute-force query every frame."""
    _print_case(f"96x96 cloth on the sphere_c4 sphere vs oracle, {cell * 100:.1f} cm cells, 6 frames")
    n, spacing, radius = 96, 0.002, 0.3
    x, f = configs.horizontal_grid(n, spacing, radius + 0.015, seed=91)
    cloth = Cloth(x, f, np.float32(1e-3), np.zeros(n * n, np.uint8))
    ox, of = make_icosphere(5, radius)
    ct = dict(thickness=0.002, stiffness=1e5, radius=0.025, cell=cell)
    params = Params(constraint_iter_count=20)
    g = GpuVBD(cloth, params)
    g.set_lanes_per_vertex(4)
    g.set_strategy(_lib.LATS_STRATEGY_GRAPH)
    g.set_obstacle(ox, of, ct["thickness"], ct["stiffness"], ct["radius"], ct["cell"])
    o = oracle_sim(cloth, params, threads=16)
    o.set_obstacle(ox, of, ct["thickness"], ct["stiffness"], ct["radius"])
    touched = 0
    full = 0
    for fr in range(6):
        g.physics_step()
        o.step()
        cg, fg = g.contacts()
        co, fo = o.contacts()
        assert np.array_equal(cg, co), fr
        assert np.array_equal(fg, fo), fr
        touched += int((cg > 0).sum())
        full += int((cg == 4).sum())
        assert np.array_equal(g.state()[0], o.state()[0].astype(np.float32)), fr
    print("contact vertex-frames", touched, "with 4 contacts", full, flush=True)
    assert touched > 0 and full > 0


def test_sphere_c4_full_size_matches_oracle(gpu_available):
    """The configured sphere drop at full size (1024x1024 cloth, 4 mm hash
    cells, 50 iterations per frame) against the oracle's query over every
    sphere face: contact sets and positions equal bit for bit over the first
    ten frames, the last three of them with thousands of vertices in contact
    (the oracle's exact prefilter keeps its brute-force query affordable at a
    million vertices)."""
    _print_case("sphere_c4 1024x1024 vs oracle, 10 frames")
    cloth, params, (ox, of), ct = configs.sphere_scene()
    g = GpuVBD(cloth, params)
    g.set_obstacle(ox, of, ct["thickness"], ct["stiffness"], ct["radius"], ct["cell"])
    o = oracle_sim(cloth, params, threads=16)
    o.set_obstacle(ox, of, ct["thickness"], ct["stiffness"], ct["radius"])
    touched = 0
    for fr in range(10):
        g.physics_step()
        o.step()
        cg, fg = g.contacts()
        co, fo = o.contacts()
        assert np.array_equal(cg, co), fr
        assert np.array_equal(fg, fo), fr
        touched += int((cg > 0).sum())
        assert np.array_equal(g.state()[0], o.state()[0].astype(np.float32)), fr
        print(f"frame {fr}: {int((cg > 0).sum())} vertices in contact", flush=True)
    assert touched > 10000
