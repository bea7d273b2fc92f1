"""Asynchronous iteration across two processors (the reference's
example2/example3 async frame): HEFT + async post-pass schedule, launch
events and execution order.

The schedule is pinned against the two Gantt dumps the reference printed for
its example2 DAG (documents/example2_scheduling_result.py:309-322, fixture
tests/golden/example2_schedules.json): processor, start and end of every
task, the position buffer it iterates, its cross-processor inputs and its
same-processor "left" link, all of which must match exactly.
"""
import ast
import json
import os
import subprocess

import numpy as np
import pytest

from tests.helpers import GOLDEN

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "bin", "test_async_schedule")


def _schedule(iters, clusters):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} not built (make -C libatssim_amd/csrc)")
    r = subprocess.run([BIN, str(iters), str(clusters)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "PASS" in r.stdout, r.stdout + r.stderr
    body = r.stdout[r.stdout.index("GRAPH_BEGIN\n") + 12:r.stdout.index("GRAPH_END")]
    lists = [ast.literal_eval("[" + line.strip()[1:-2] + "]") for line in body.strip().split("\n")]
    events = {}
    for line in r.stdout.split("\n"):
        if line.startswith("EVENTS"):
            p, rest = line[7:].split(":", 1)
            events[int(p)] = rest.strip()
    order = [int(v) for v in r.stdout.split("ORDER:")[1].split("\n")[0].split()]
    return lists, events, order


def _golden(iters):
    return json.load(open(os.path.join(GOLDEN, "example2_schedules.json")))[f"iters{iters}"]


@pytest.mark.parametrize("iters", [1, 10])
def test_schedule_matches_reference_dump(iters):
    (cpu, gpu, ins, left), _, _ = _schedule(iters, 10)
    g = _golden(iters)
    for name, ours, ref in (("cpu", cpu, g["cpu"]), ("gpu", gpu, g["gpu"]), ("ins", ins, g["ins"]),
                            ("left", left, g["left"])):
        assert len(ours) == len(ref), (name, len(ours), len(ref))
        for a, b in zip(ours, ref):
            # integers exact, times as printed (3 decimals)
            assert list(a[:-2]) == list(b[:-2]), (name, a, b)
            assert abs(a[-2] - b[-2]) < 1e-9 and abs(a[-1] - b[-1]) < 1e-9, (name, a, b)


def test_iter1_launch_events_and_order():
    """Wait/signal events of the 1-iteration schedule (launcher.cpp:3364-3714)
    and the LaunchModeSequeceHetero order (launcher.cpp:1383-1473): CPU event
    k first runs every GPU event up to the one it waits for."""
    (cpu, gpu, _, _), events, order = _schedule(1, 10)
    # CPU: root, predict | 1 | 2 | 4 | 6 (waits GPU 0) | 8 (GPU 1) | 10 (GPU 2) | last, update, terminal (GPU 3)
    assert events[0] == "[0-1 w-1 s0] [2-2 w-1 s1] [3-3 w-1 s2] [4-4 w-1 s3] [5-5 w0 s-1] [6-6 w1 s-1] " \
                        "[7-7 w2 s-1] [8-10 w3 s-1]"
    assert events[1] == "[0-0 w0 s0] [1-1 w1 s1] [2-2 w2 s2] [3-3 w3 s3]"
    assert order == [0, 1, 2, 4, 3, 6, 5, 8, 7, 10, 9, 11, 12]


@pytest.mark.parametrize("iters,clusters", [(3, 4), (20, 10), (100, 10), (5, 1)])
def test_async_schedules_run_in_dependency_order(iters, clusters):
    """Every task runs once, after its predecessors (the test binary checks),
    and every iterative task has a buffer."""
    (cpu, gpu, ins, left), _, order = _schedule(iters, clusters)
    for t in cpu + gpu:
        colour, tid, buf = t[0], t[1], t[2]
        if colour >= 14:               # vbd_all_in_one tasks
            assert buf != 4294967295, t


# ---------------------------------------------------------------------------
# the asynchronous frame on the oracle vs the reference's published history
# ---------------------------------------------------------------------------
def launch_sequence(iters, clusters):
    """The LaunchModeSequeceHetero sequence of the async schedule: a list of
    (proc, fn, iter, cluster, buffer, left, right, main, ins)."""
    r = subprocess.run([BIN, str(iters), str(clusters)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "PASS" in r.stdout, r.stdout + r.stderr
    seq = []
    for line in r.stdout.split("\n"):
        if not line.startswith("PARAM "):
            continue
        v = [int(t) for t in line.split()[1:]]
        k, proc, fn, it, cl, buf, left, right, main, nin = v[:10]
        ins = v[10:10 + nin]
        if fn >= 0:
            seq.append((proc, fn, it, cl, buf, left, right, main, tuple(ins)))
    return seq


def replay_on_oracle(sim, seq, record):
    sim.async_begin()
    for proc, fn, it, cl, buf, left, right, main, ins in seq:
        sim.dispatch(fn, it, cl, buf, left, right, main, ins, record=record)
    return sim.energies()


def test_square8k_async_frame_matches_published():
    """example2's main (examples/example2_async_iter.cpp:1752-1805): 30 frames,
    then one asynchronous frame from that state with the energy recorded at
    the reference's hook points: 102 values, published as list_async
    (documents/example2_async_convergence.py:82-85).  The oracle replays the
    async schedule task by task (CpuSolver::fn_dispatch with its 32+32
    position buffers and read_and_solve_conflict merges)."""
    from oracle import oracle as O
    from libatssim_amd import configs
    from tests.helpers import square8k_mesh
    x, f = square8k_mesh()
    cloth = configs.square8k_cloth(x, f)
    p = configs.square8k_params()
    m = O.Mesh(x, f)
    assert m.nclusters == 10
    seq = launch_sequence(100, 10)
    s = O.Sim(m, cloth.mass, cloth.fixed, implicit_dt=p.implicit_dt, iter_count=100,
              k_stretch=p.stiffness_stretch_spring, k_bending=p.stiffness_quadratic_bending,
              damping=p.damping_cloth, threads=O.max_threads())
    for _ in range(30):
        s.step()
    e = replay_on_oracle(s, seq, record=True)
    pub = np.array(json.load(open(os.path.join(GOLDEN, "square8k_published.json")))["list_async"])
    assert len(e) == len(pub) == 102
    rel = np.abs(e - pub) / np.abs(pub)
    assert rel.max() < 5e-6, (rel.max(), np.argmax(rel), e[:5], pub[:5])


def test_two_memory_harness_without_device():
    """The heterogeneous-launch harness with both processors on oracles
    (tests/cpp/test_hetero_vbd.cpp, "cpu" mode): runs, and the two-memory
    replay stays close to the shared-buffer one."""
    path = os.path.join(os.path.dirname(BIN), "test_hetero_vbd")
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (make -C libatssim_amd/csrc)")
    r = subprocess.run([path, "24", "20", "2", "cpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "PASS" in r.stdout, r.stdout + r.stderr
