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
