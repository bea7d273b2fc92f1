"""lats_vbd_energy_noalloc (the entry point that allocates nothing) against
lats_vbd_energy_now (the allocating entry point with the same kernels):
bit-identical energies on two calls in a row, the second on a smaller cloth,
so the fixed per-GPU partials array is reused with fewer blocks; the
positions argument (NULL or an explicit device buffer); the size limit."""
import ctypes

import numpy as np
import pytest

from libatssim_amd import GpuVBD, Params, _lib, configs
from libatssim_amd._lib import lib

pytestmark = pytest.mark.gpu


class _Hip:
    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p

    def read_f32(self, p, n):
        out = np.zeros(n, np.float32)
        assert self.h.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, 4 * n, 2) == 0   # device to host
        return out

    def free(self, p):
        assert self.h.hipFree(p) == 0


def _windy_context(n, seed, frames):
    cloth = configs.grid_cloth(n, "top", seed=seed)
    g = GpuVBD(cloth, Params(constraint_iter_count=20))
    g.set_external_force([0.3, 0.0, -0.2])
    for _ in range(frames):
        g.physics_step()
    g.frame_begin()          # a running frame: predicted positions, then a few sweeps
    g.iterate(3)
    g.synchronize()
    return g


def _noalloc(g, x_dev, out):
    _lib.check(lib().lats_vbd_energy_noalloc(g.handle, x_dev, out))
    g.synchronize()


def test_noalloc_energy_matches_energy_now(gpu_available):
    hip = _Hip()
    out = hip.malloc(64)
    try:
        # first call: 96 x 96 cloth; second call: 40 x 40 cloth (fewer partials)
        for n, seed in ((96, 11), (40, 12)):
            g = _windy_context(n, seed, 3)
            try:
                for call in range(2):
                    _noalloc(g, None, out)
                    e_na = hip.read_f32(out, 1)[0]
                    e_ref = g.energy_now()
                    print(f"cloth {n}x{n} call {call}: noalloc {e_na!r} energy_now {e_ref!r}", flush=True)
                    assert np.float32(e_na).tobytes() == np.float32(e_ref).tobytes()
                    # the same positions passed explicitly
                    xd = ctypes.c_void_p()
                    _lib.check(lib().lats_vbd_device_positions(g.handle, ctypes.byref(xd)))
                    _noalloc(g, xd, out)
                    assert np.float32(hip.read_f32(out, 1)[0]).tobytes() == np.float32(e_ref).tobytes()
                    g.iterate(2)
                    g.synchronize()
            finally:
                g.close()
    finally:
        hip.free(out)


def test_noalloc_energy_rejects_oversized_cloth(gpu_available):
    # 1100 x 1100: 1,210,000 vertices + 3,625,601 springs + 3,621,205 pairs > 8,388,608
    cloth = configs.grid_cloth(1100, "top", seed=13)
    g = GpuVBD(cloth, Params(constraint_iter_count=1))
    hip = _Hip()
    out = hip.malloc(64)
    try:
        rc = lib().lats_vbd_energy_noalloc(g.handle, None, out)
        assert rc == _lib.LATS_ERR_INVALID_ARGUMENT
        assert b"too large" in lib().lats_last_error()
        # the allocating entry point still works on it
        assert np.isfinite(g.energy_now())
    finally:
        hip.free(out)
        g.close()


def test_noalloc_energy_null_output(gpu_available):
    g = _windy_context(16, 14, 1)
    try:
        assert lib().lats_vbd_energy_noalloc(g.handle, None, None) == _lib.LATS_ERR_INVALID_ARGUMENT
    finally:
        g.close()
