"""Launcher (task registration + HEFT scheduler) tests, written as C++
programs like the reference's examples and run here."""
import os
import subprocess

import pytest

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "bin")


def _run(name, timeout=120):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (make -C libatssim_amd/csrc)")
    r = subprocess.run([path], capture_output=True, text=True, timeout=timeout)
    print(r.stdout[-2000:])
    return r


def test_heft_2002_known_answer():
    """Topcuoglu DAG: makespan 91, speedups 39.56/42.86/57.14 % (README.md:25-31)."""
    r = _run("test_heft_2002")
    assert r.returncode == 0 and "PASS" in r.stdout
    assert "makespan 91" in r.stdout


@pytest.mark.gpu
def test_vbd_dag_through_scheduler_matches_fused(gpu_available):
    print("[case] VBD DAG via Launcher LaunchModeGpu", flush=True)
    r = _run("test_vbd_dag")
    assert r.returncode == 0 and "PASS" in r.stdout, r.stdout + r.stderr


@pytest.mark.gpu
def test_integration_snippets_match_fused_frames(gpu_available):
    """The adapter, the hand-written LaunchParam slot and the XPBD DAG of the
    integration guide, run for two frames each, equal the fused frames."""
    print("[case] integration snippets via Launcher LaunchModeGpu", flush=True)
    r = _run("test_integration_snippets")
    assert r.returncode == 0 and "PASS" in r.stdout, r.stdout + r.stderr


def test_hetero_launch_stops_after_failed_host_wait():
    """A host wait for the queue that fails (a time-out) ends the host side of
    LaunchModeHetero / ProgressiveHetero: no further host event runs and the
    queue is released and drained (host-only hooks)."""
    r = _run("test_hetero_wait_failure")
    assert r.returncode == 0 and "PASS" in r.stdout, r.stdout + r.stderr


@pytest.mark.gpu
def test_waiting_events_at_depth(gpu_available):
    """60 alternating host / GPU events x 100 loops on one shared counter
    (the reference's test_1_waiting_events): every loop ends at 120 and no GPU
    event completes before the host signal it waits for."""
    print("[case] host/GPU waiting events, depth 60 x 100 loops", flush=True)
    r = _run("test_waiting_events", timeout=300)
    assert r.returncode == 0 and "PASS" in r.stdout, r.stdout + r.stderr


def test_patch_layout_bounds_and_gauss_seidel_order():
    """Patch strategy layout (CPU): every index a patch launch computes is in
    bounds for 4/8/16 lanes and 256/512/1024 threads, every patch and slot is
    covered once per launch, and emulated launches reproduce the plain
    colour-by-colour sweep bit for bit (grids of even, odd and narrow sizes,
    a mesh with holes, the 256x256 cloth)."""
    r = _run("test_patch_layout")
    assert r.returncode == 0 and "PASS" in r.stdout, r.stdout + r.stderr
