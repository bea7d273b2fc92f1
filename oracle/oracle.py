"""ctypes front-end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, by the smoke() entry point (to check its result) and by
bench.py's cpu_baseline leg.  The product path never imports this module.

The oracle is a C++ restatement of the reference's VBD cloth path
(vbd_oracle.cpp in this directory cites the reference file:line of every
function).  Build it with `make -C oracle`.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_u8p = ctypes.POINTER(ctypes.c_uint8)


def _ptr(a, t):
    return a.ctypes.data_as(t)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        L.orc_mesh_create.restype = ctypes.c_void_p
        L.orc_mesh_create.argtypes = [ctypes.c_uint32, _f32p, ctypes.c_uint32, _i32p]
        L.orc_mesh_destroy.argtypes = [ctypes.c_void_p]
        L.orc_mesh_counts.argtypes = [ctypes.c_void_p, _u32p]
        L.orc_mesh_edges.argtypes = [ctypes.c_void_p, _i32p]
        L.orc_mesh_bending.argtypes = [ctypes.c_void_p, _i32p]
        L.orc_mesh_csr_size.restype = ctypes.c_uint32
        L.orc_mesh_csr_size.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_mesh_csr.argtypes = [ctypes.c_void_p, ctypes.c_int, _u32p]
        L.orc_mesh_rest.argtypes = [ctypes.c_void_p, _f32p, _f32p]
        L.orc_sim_create.restype = ctypes.c_void_p
        L.orc_sim_create.argtypes = [ctypes.c_void_p, _f32p, _u32p, _f32p, _u8p, ctypes.c_int]
        L.orc_sim_destroy.argtypes = [ctypes.c_void_p]
        L.orc_sim_set_threads.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_sim_set_energy_bending.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_sim_set_wind.argtypes = [ctypes.c_void_p, _f32p]
        L.orc_sim_step.restype = ctypes.c_uint32
        L.orc_sim_step.argtypes = [ctypes.c_void_p, ctypes.c_int, _f32p, ctypes.c_uint32]
        L.orc_sim_get_state.argtypes = [ctypes.c_void_p, _f64p, _f64p]
        L.orc_sim_set_state.argtypes = [ctypes.c_void_p, _f64p, _f64p]
        L.orc_sim_energy.restype = ctypes.c_double
        L.orc_sim_energy.argtypes = [ctypes.c_void_p]
        L.orc_sim_set_positions.argtypes = [ctypes.c_void_p, _f64p]
        L.orc_sim_forces.argtypes = [ctypes.c_void_p, _f64p]
        L.orc_max_threads.restype = ctypes.c_int
        L.orc_sim_set_obstacle.argtypes = [ctypes.c_void_p, ctypes.c_uint32, _f32p, ctypes.c_uint32, _i32p,
                                           ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int]
        L.orc_sim_set_contact_prefilter.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_sim_contacts.restype = ctypes.c_uint32
        L.orc_sim_contacts.argtypes = [ctypes.c_void_p, _u32p, _i32p, ctypes.c_uint32]
        L.orc_sim_async_begin.argtypes = [ctypes.c_void_p]
        L.orc_sim_step_xpbd.restype = ctypes.c_uint32
        L.orc_sim_step_xpbd.argtypes = [ctypes.c_void_p, ctypes.c_int, _f32p, ctypes.c_uint32]
        L.orc_sim_xpbd_clusters.restype = ctypes.c_uint32
        L.orc_sim_xpbd_clusters.argtypes = [ctypes.c_void_p, ctypes.c_int, _u32p, ctypes.c_uint32]
        L.orc_sim_dispatch.argtypes = [ctypes.c_void_p] + [ctypes.c_uint32] * 7 + [_u32p, ctypes.c_int, ctypes.c_int]
        L.orc_sim_energies.restype = ctypes.c_uint32
        L.orc_sim_energies.argtypes = [ctypes.c_void_p, _f32p, ctypes.c_uint32]
        L.orc_template_vf.restype = ctypes.c_int
        L.orc_template_vf.argtypes = [_f32p, ctypes.c_float, _f32p, _f32p]
        _LIB = L
    return _LIB


CSR_ADJ_EDGES, CSR_ADJ_BENDING, CSR_ADJ_VERTS, CSR_ADJ_VERTS_BENDING, CSR_CLUSTERS = range(5)


class Mesh:
    """Topology + colouring built by the oracle from positions and faces."""

    def __init__(self, x, faces):
        self.x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 3)
        self.faces = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 3)
        self._h = lib().orc_mesh_create(len(self.x), _ptr(self.x, _f32p), len(self.faces), _ptr(self.faces, _i32p))
        c = np.zeros(4, np.uint32)
        lib().orc_mesh_counts(self._h, _ptr(c, _u32p))
        self.nv, self.ne, self.nb, self.nclusters = (int(v) for v in c)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_mesh_destroy(self._h)
            self._h = None

    def edges(self):
        out = np.zeros((self.ne, 2), np.int32)
        lib().orc_mesh_edges(self._h, _ptr(out, _i32p))
        return out

    def bending(self):
        out = np.zeros((self.nb, 4), np.int32)
        lib().orc_mesh_bending(self._h, _ptr(out, _i32p))
        return out

    def csr(self, which):
        n = lib().orc_mesh_csr_size(self._h, which)
        out = np.zeros(n, np.uint32)
        lib().orc_mesh_csr(self._h, which, _ptr(out, _u32p))
        return out

    def clusters(self):
        """List of vertex-id arrays, one per colour (reference order)."""
        p = self.csr(CSR_CLUSTERS)
        n = self.nclusters
        return [p[p[i]:p[i + 1]] for i in range(n)]

    def rest(self):
        rl = np.zeros(self.ne, np.float32)
        q = np.zeros((self.nb, 4, 4), np.float32)   # q[b, col, row]
        lib().orc_mesh_rest(self._h, _ptr(rl, _f32p), _ptr(q, _f32p))
        return rl, q


class Sim:
    """One cloth stepped by the oracle's VBD (FP32 or FP64)."""

    def __init__(self, mesh, mass, fixed, implicit_dt=1.0 / 60.0, num_substep=1, iter_count=100,
                 k_stretch=1e4, k_bending=1e-2, damping=2.0, gravity=(0.0, -9.8, 0.0), fp64=False,
                 energy_bending=False, threads=1):
        self.mesh = mesh
        p = np.array([implicit_dt, k_stretch, k_bending, damping, *gravity, 0, 0], np.float32)
        ip = np.array([num_substep, iter_count], np.uint32)
        self.mass = np.ascontiguousarray(mass, np.float32)
        self.fixed = np.ascontiguousarray(fixed, np.uint8)
        self._h = lib().orc_sim_create(mesh._h, _ptr(p, _f32p), _ptr(ip, _u32p), _ptr(self.mass, _f32p),
                                       _ptr(self.fixed, _u8p), int(fp64))
        self.iter_count = iter_count
        lib().orc_sim_set_energy_bending(self._h, int(energy_bending))
        lib().orc_sim_set_threads(self._h, int(threads))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_sim_destroy(self._h)
            self._h = None

    def set_threads(self, n):
        lib().orc_sim_set_threads(self._h, int(n))

    def set_wind(self, w):
        w = np.ascontiguousarray(w, np.float32)
        lib().orc_sim_set_wind(self._h, _ptr(w, _f32p))

    def step(self, record=False):
        cap = self.iter_count + 1
        e = np.zeros(cap, np.float32)
        n = lib().orc_sim_step(self._h, int(record), _ptr(e, _f32p), cap)
        return e[:n]

    def state(self):
        x = np.zeros((self.mesh.nv, 3), np.float64)
        v = np.zeros((self.mesh.nv, 3), np.float64)
        lib().orc_sim_get_state(self._h, _ptr(x, _f64p), _ptr(v, _f64p))
        return x, v

    def set_state(self, x, v):
        x = np.ascontiguousarray(x, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        lib().orc_sim_set_state(self._h, _ptr(x, _f64p), _ptr(v, _f64p))

    def set_positions(self, x):
        """Positions alone; the frame start and velocities stay."""
        x = np.ascontiguousarray(x, np.float64)
        lib().orc_sim_set_positions(self._h, _ptr(x, _f64p))

    def forces(self):
        """Per-vertex force vector of the vertex solve at the current positions."""
        f = np.zeros(self.mesh.nv * 3, np.float64)
        lib().orc_sim_forces(self._h, _ptr(f, _f64p))
        return f.reshape(-1, 3)

    def energy(self):
        return lib().orc_sim_energy(self._h)

    # XPBD (CpuSolver::physics_step_xpbd)
    def step_xpbd(self, record=False):
        cap = 2 * self.iter_count + 2
        e = np.zeros(cap, np.float32)
        n = lib().orc_sim_step_xpbd(self._h, int(record), _ptr(e, _f32p), cap)
        return e[:n]

    def xpbd_clusters(self, which):
        n = lib().orc_sim_xpbd_clusters(self._h, int(which), None, 0)
        p = np.zeros(n, np.uint32)
        lib().orc_sim_xpbd_clusters(self._h, int(which), _ptr(p, _u32p), n)
        k = int(p[0]) - 1
        return [p[p[i]:p[i + 1]] for i in range(k)]

    # asynchronous frame (CpuSolver::physics_step_vbd_async + fn_dispatch)
    def async_begin(self):
        lib().orc_sim_async_begin(self._h)

    def dispatch(self, fn, iter_idx, cluster_idx, buffer_idx, left, right, main, ins=(), record=False):
        ins = np.ascontiguousarray(ins, np.uint32)
        lib().orc_sim_dispatch(self._h, int(fn), int(iter_idx), int(cluster_idx), int(buffer_idx), int(left),
                               int(right), len(ins), _ptr(ins, _u32p) if len(ins) else None, int(main), int(record))

    def energies(self):
        n = lib().orc_sim_energies(self._h, None, 0)
        e = np.zeros(max(n, 1), np.float32)
        lib().orc_sim_energies(self._h, _ptr(e, _f32p), n)
        return e[:n]

    def set_obstacle(self, x, faces, thickness, stiffness, radius, max_contacts=4):
        self._obs = (np.ascontiguousarray(x, np.float32), np.ascontiguousarray(faces, np.int32))
        x, f = self._obs
        lib().orc_sim_set_obstacle(self._h, len(x), _ptr(x, _f32p), len(f), _ptr(f, _i32p), float(thickness),
                                   float(stiffness), float(radius), int(max_contacts))

    def set_contact_prefilter(self, on):
        """The exact prefilter of the contact query (default on; off = every face)."""
        lib().orc_sim_set_contact_prefilter(self._h, int(bool(on)))

    def contacts(self, max_contacts=4):
        """(counts[V], face ids[V, max_contacts] ascending, -1 padded)."""
        n = self.mesh.nv
        counts = np.zeros(n, np.uint32)
        tot = lib().orc_sim_contacts(self._h, _ptr(counts, _u32p), None, 0)
        flat = np.zeros(max(tot, 1), np.int32)
        lib().orc_sim_contacts(self._h, _ptr(counts, _u32p), _ptr(flat, _i32p), tot)
        out = np.full((n, max_contacts), -1, np.int32)
        pos = 0
        for v in range(n):
            k = int(counts[v])
            out[v, :k] = flat[pos:pos + k]
            pos += k
        return counts, out


def template_vf(p, a, b, c, min_distance):
    """The reference's vertex-face narrow phase: (hit, distance, barycentric)."""
    pts = np.ascontiguousarray(np.stack([p, a, b, c]), np.float32)
    d = np.zeros(1, np.float32)
    bary = np.zeros(3, np.float32)
    hit = lib().orc_template_vf(_ptr(pts, _f32p), float(min_distance), _ptr(d, _f32p), _ptr(bary, _f32p))
    return bool(hit), float(d[0]), bary


def max_threads():
    return int(lib().orc_max_threads())
