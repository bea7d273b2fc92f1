"""Latency-floor measurement (lats_time_floor): every kind returns a
positive time, a gather costs at least an empty launch, and bad arguments
are rejected before anything runs."""
import ctypes

import pytest

from libatssim_amd import _lib

pytestmark = pytest.mark.gpu


def _floor(kind, threads, launches=50):
    us = ctypes.c_float()
    _lib.check(_lib.measure_lib().lats_time_floor(kind, threads, launches, ctypes.byref(us)))
    return us.value


def test_floor_kinds(gpu_available):
    print("[case] launch floors, 50 launches", flush=True)
    empty = _floor(_lib.LATS_FLOOR_EMPTY, 65536)
    gather = _floor(_lib.LATS_FLOOR_GATHER, 65536)
    barrier = _floor(_lib.LATS_FLOOR_GRID_BARRIER, 64 * 256)
    reread = _floor(_lib.LATS_FLOOR_REREAD, 131072)
    for kind in (_lib.LATS_FLOOR_HANDOFF_FENCED, _lib.LATS_FLOOR_HANDOFF_RELAXED, _lib.LATS_FLOOR_HANDOFF_XCD):
        assert 0.0 < _floor(kind, 64 * 256) < 10000.0
    assert 0.0 < empty < 1000.0
    assert 0.0 < reread < 1000.0
    assert gather >= 0.5 * empty
    assert 0.0 < barrier < 10000.0


def test_floor_rejects(gpu_available):
    us = ctypes.c_float()
    assert _lib.measure_lib().lats_time_floor(7, 64, 10, ctypes.byref(us)) == 1
    assert _lib.measure_lib().lats_time_floor(-1, 64, 10, ctypes.byref(us)) == 1
    assert _lib.measure_lib().lats_time_floor(_lib.LATS_FLOOR_EMPTY, 0, 10, ctypes.byref(us)) == 1
    assert _lib.measure_lib().lats_time_floor(_lib.LATS_FLOOR_GRID_BARRIER, 1 << 24, 10, ctypes.byref(us)) == 1
