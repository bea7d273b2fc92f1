"""Host-side Python front-end of the VBD cloth GPU slot.

Mirrors the reference's driver interface for the VBD path:
`SolverInterface` (examples/example2_async_iter.cpp:1550-1710: physics_step,
restart_system, save_current_frame_state, load_saved_state,
save_mesh_to_obj) over the HIP context of include/libatssim_hip.h.  The
topology, colouring and clusters come from the library's own builder
(lats_topology_*), in the reference's formats.
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import check, lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------------------
# scenes
# ---------------------------------------------------------------------------
def make_grid(nx, ny, spacing=0.01, jitter=1e-4, seed=0):
    """Seeded nx x ny grid cloth in the xy plane (row 0 on top, y = 0)."""
    x = np.zeros((nx * ny, 3), np.float32)
    f = np.zeros(((nx - 1) * (ny - 1) * 2, 3), np.int32)
    check(lib().lats_scene_grid(nx, ny, spacing, jitter, seed, _p(x), _p(f)))
    return x, f


def _obj_arrays(h):
    counts = np.zeros(3, np.uint32)
    try:
        check(lib().lats_obj_counts(h, _p(counts)))
        x = np.zeros((int(counts[0]), 3), np.float32)
        f = np.zeros((int(counts[1]), 3), np.int32)
        check(lib().lats_obj_copy(h, _p(x), _p(f)))
    finally:
        lib().lats_obj_destroy(h)
    return x, f, int(counts[2])


def read_obj(path):
    """(positions float32 [V,3], triangles int32 [F,3], dropped) of an OBJ file,
    read the way the reference's mesh reader does (fan triangulation, faces
    repeating a vertex dropped)."""
    h = ctypes.c_void_p()
    check(lib().lats_obj_read(os.fsencode(path), ctypes.byref(h)))
    return _obj_arrays(h)


def parse_obj(text):
    h = ctypes.c_void_p()
    check(lib().lats_obj_parse(text.encode(), ctypes.byref(h)))
    return _obj_arrays(h)


def write_obj(path, x, faces, mesh_id=0):
    """One frame as OBJ (the reference's save_mesh_to_obj format)."""
    x = np.ascontiguousarray(x, np.float32)
    f = np.ascontiguousarray(faces, np.int32)
    check(lib().lats_obj_write(os.fsencode(path), _p(x), len(x), _p(f), len(f), int(mesh_id)))


def make_icosphere(level, radius, centre=(0.0, 0.0, 0.0)):
    """Icosahedron subdivided `level` times on a sphere; faces point outwards."""
    c = np.ascontiguousarray(centre, np.float32)
    nv = ctypes.c_uint32()
    nf = ctypes.c_uint32()
    check(lib().lats_scene_icosphere(level, radius, _p(c), None, None, ctypes.byref(nv), ctypes.byref(nf)))
    x = np.zeros((nv.value, 3), np.float32)
    f = np.zeros((nf.value, 3), np.int32)
    check(lib().lats_scene_icosphere(level, radius, _p(c), _p(x), _p(f), ctypes.byref(nv), ctypes.byref(nf)))
    return x, f


def make_wind(num_frames, num_verts, max_total=2.0, seed=0):
    w = np.zeros((num_frames, 3), np.float32)
    check(lib().lats_scene_wind(num_frames, num_verts, max_total, seed, _p(w)))
    return w


# ---------------------------------------------------------------------------
# topology
# ---------------------------------------------------------------------------
class Topology:
    """Edges, bending pairs, rest state, adjacency CSRs and colour clusters."""

    _ARRAYS = {
        "edges": (_lib.LATS_TOPO_EDGES, np.int32, 2),
        "bending": (_lib.LATS_TOPO_BENDING, np.int32, 4),
        "rest_length": (_lib.LATS_TOPO_REST_LENGTH, np.float32, 1),
        "bending_Q": (_lib.LATS_TOPO_BENDING_Q, np.float32, 16),
        "adj_edges_csr": (_lib.LATS_TOPO_ADJ_EDGES_CSR, np.uint32, 1),
        "adj_bending_csr": (_lib.LATS_TOPO_ADJ_BENDING_CSR, np.uint32, 1),
        "clusters_csr": (_lib.LATS_TOPO_CLUSTERS_CSR, np.uint32, 1),
        "colour": (_lib.LATS_TOPO_COLOUR, np.uint8, 1),
        "faces": (_lib.LATS_TOPO_FACES, np.int32, 3),
    }
    # file name of each array on disk (lats_topology_save / _load)
    _FILES = {"faces": "faces", "edges": "edges", "bending": "bending_edges", "rest_length": "edges_rest_state_length",
              "bending_Q": "bending_edges_Q", "adj_edges_csr": "vert_adj_edges",
              "adj_bending_csr": "vert_adj_bending_edges", "clusters_csr": "clusterd_per_vertex_bending",
              "colour": "per_vertex_bending_cluster_id"}

    def __init__(self, x, faces):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 3)
        faces = np.ascontiguousarray(faces, np.int32).reshape(-1, 3)
        h = ctypes.c_void_p()
        check(lib().lats_topology_build(_p(x), len(x), _p(faces), len(faces), ctypes.byref(h)))
        self._take(h)

    @classmethod
    def load(cls, directory):
        """The topology saved in `directory` (lats_topology_load, checked)."""
        h = ctypes.c_void_p()
        check(lib().lats_topology_load(os.fsencode(directory), ctypes.byref(h)))
        t = cls.__new__(cls)
        t._take(h)
        return t

    def save(self, directory):
        """One raw array per file, the format lats_topology_save writes."""
        for name, fname in self._FILES.items():
            np.ascontiguousarray(getattr(self, name)).tofile(os.path.join(directory, fname + ".dat"))

    def _take(self, h):
        """Copies every array out of the library handle h and frees it."""
        try:
            c = np.zeros(4, np.uint32)
            check(lib().lats_topology_counts(h, _p(c)))
            self.num_verts, self.num_edges, self.num_bending, self.num_clusters = (int(v) for v in c)
            for name, (which, dt, w) in self._ARRAYS.items():
                n = ctypes.c_uint64()
                check(lib().lats_topology_size(h, which, ctypes.byref(n)))
                a = np.zeros(n.value, dt)
                if n.value:
                    check(lib().lats_topology_copy(h, which, _p(a)))
                setattr(self, name, a.reshape(-1, w) if w > 1 else a)
        finally:
            lib().lats_topology_destroy(h)
        self.bending_Q = self.bending_Q.reshape(-1, 4, 4)

    def clusters(self):
        p = self.clusters_csr
        return [p[p[i]:p[i + 1]] for i in range(self.num_clusters)]


# ---------------------------------------------------------------------------
# parameters of a run
# ---------------------------------------------------------------------------
class Params:
    """The SceneParams fields the VBD path reads (scene_params.h:9-163)."""

    def __init__(self, implicit_dt=1.0 / 60.0, num_substep=1, constraint_iter_count=100,
                 stiffness_stretch_spring=1e4, stiffness_quadratic_bending=1e-2, damping_cloth=2.0,
                 gravity=(0.0, -9.8, 0.0)):
        self.implicit_dt = float(np.float32(implicit_dt))
        self.num_substep = int(num_substep)
        self.constraint_iter_count = int(constraint_iter_count)
        self.stiffness_stretch_spring = float(stiffness_stretch_spring)
        self.stiffness_quadratic_bending = float(stiffness_quadratic_bending)
        self.damping_cloth = float(damping_cloth)
        self.gravity = tuple(float(g) for g in gravity)

    def c_struct(self):
        s = _lib.SceneParams()
        s.implicit_dt = self.implicit_dt
        s.num_substep = self.num_substep
        s.constraint_iter_count = self.constraint_iter_count
        s.stiffness_stretch_spring = self.stiffness_stretch_spring
        s.stiffness_quadratic_bending = self.stiffness_quadratic_bending
        s.damping_cloth = self.damping_cloth
        s.gravity = (ctypes.c_float * 3)(*self.gravity)
        return s


class Cloth:
    """One cloth: rest positions, faces, per-vertex mass and pins, topology."""

    def __init__(self, x, faces, mass, fixed, topology=None):
        self.x = np.ascontiguousarray(x, np.float32).reshape(-1, 3)
        self.faces = np.ascontiguousarray(faces, np.int32).reshape(-1, 3)
        nv = len(self.x)
        self.mass = np.ascontiguousarray(np.broadcast_to(np.asarray(mass, np.float32), (nv,)))
        self.fixed = np.ascontiguousarray(np.asarray(fixed).astype(np.uint8).reshape(nv))
        self.topology = topology if topology is not None else Topology(self.x, self.faces)

    @property
    def num_verts(self):
        return len(self.x)


# ---------------------------------------------------------------------------
# GPU solver
# ---------------------------------------------------------------------------
class GpuVBD:
    """Single-cloth VBD on the GPU (the reference's GpuSolver VBD path)."""

    def __init__(self, cloth, params, strategy=_lib.LATS_STRATEGY_AUTO):
        self.cloth = cloth
        self.params = params
        t = cloth.topology
        self._keep = [cloth.x, t.edges, t.bending, t.rest_length, t.bending_Q, t.adj_edges_csr,
                      t.adj_bending_csr, t.clusters_csr, cloth.mass, cloth.fixed]
        d = _lib.MeshDesc()
        d.num_verts = t.num_verts
        d.num_edges = t.num_edges
        d.num_bending_edges = t.num_bending
        d.num_clusters = t.num_clusters
        d.rest_x = _p(cloth.x).value
        d.edges = _p(t.edges).value
        d.bending_edges = _p(t.bending).value
        d.edges_rest_length = _p(t.rest_length).value
        d.bending_Q = _p(t.bending_Q).value
        d.vert_adj_edges_csr = _p(t.adj_edges_csr).value
        d.vert_adj_bending_csr = _p(t.adj_bending_csr).value
        d.clusters_csr = _p(t.clusters_csr).value
        d.vert_mass = _p(cloth.mass).value
        d.is_fixed = _p(cloth.fixed).value
        self._params = params.c_struct()
        h = ctypes.c_void_p()
        check(lib().lats_vbd_create(ctypes.byref(d), ctypes.byref(self._params), ctypes.byref(h)))
        self._h = h
        _lib.track(self)
        if strategy != _lib.LATS_STRATEGY_AUTO:
            check(lib().lats_vbd_set_strategy(self._h, strategy))

    def close(self):
        if getattr(self, "_h", None):
            lib().lats_vbd_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def handle(self):
        return self._h

    def set_strategy(self, strategy, persistent_blocks=None):
        check(lib().lats_vbd_set_strategy(self._h, int(strategy)))
        if persistent_blocks is not None:
            check(lib().lats_vbd_set_persistent_blocks(self._h, int(persistent_blocks)))

    def set_solver(self, solver):
        """"vbd" (default) or "xpbd" (the reference's XPBD path, same context)."""
        code = {"vbd": _lib.LATS_SOLVER_VBD, "xpbd": _lib.LATS_SOLVER_XPBD}.get(solver, solver)
        check(lib().lats_vbd_set_solver(self._h, int(code)))

    def solver(self):
        v = ctypes.c_int()
        check(lib().lats_vbd_get_solver(self._h, ctypes.byref(v)))
        return {_lib.LATS_SOLVER_VBD: "vbd", _lib.LATS_SOLVER_XPBD: "xpbd"}[v.value]

    def xpbd_clusters(self, which):
        """XPBD constraint clusters (which: 0 springs, 1 bending pairs): list of id arrays."""
        n = ctypes.c_uint64()
        check(lib().lats_vbd_xpbd_clusters(self._h, int(which), None, 0, ctypes.byref(n)))
        csr = np.zeros(n.value, np.uint32)
        check(lib().lats_vbd_xpbd_clusters(self._h, int(which), _p(csr), n.value, ctypes.byref(n)))
        k = int(csr[0]) - 1
        return [csr[csr[i]:csr[i + 1]] for i in range(k)]

    def set_regions(self, regions):
        """Bands of the colour-pass block mapping (1..8, results identical)."""
        check(lib().lats_vbd_set_regions(self._h, int(regions)))

    def set_patches(self, patches=0, colours_per_launch=0, threads=1024):
        """Patch strategy geometry: patches (multiple of 8, 0 = derived),
        colour passes per launch (0 = a whole sweep), threads per workgroup
        (256, 512, 1024).  Results are identical for every choice."""
        check(lib().lats_vbd_set_patches(self._h, int(patches), int(colours_per_launch), int(threads)))

    def patch_info(self):
        """The patch layout in use (built on first call): dict of lats_patch_info."""
        info = _lib.PatchInfo()
        check(lib().lats_vbd_get_patches(self._h, ctypes.byref(info)))
        return {name: getattr(info, name) for name, _ in info._fields_}

    def set_colour_block(self, threads):
        check(lib().lats_vbd_set_colour_block(self._h, int(threads)))

    def colour_block(self):
        t = ctypes.c_uint32()
        check(lib().lats_vbd_get_colour_block(self._h, ctypes.byref(t)))
        return t.value

    def regions(self):
        n = ctypes.c_uint32()
        check(lib().lats_vbd_get_regions(self._h, ctypes.byref(n)))
        return n.value

    def set_lanes_per_vertex(self, lanes):
        check(lib().lats_vbd_set_lanes_per_vertex(self._h, int(lanes)))

    def lanes_per_vertex(self):
        n = ctypes.c_uint32()
        check(lib().lats_vbd_get_lanes_per_vertex(self._h, ctypes.byref(n)))
        return n.value

    def strategy(self):
        st = ctypes.c_int()
        nb = ctypes.c_uint32()
        check(lib().lats_vbd_get_strategy(self._h, ctypes.byref(st), ctypes.byref(nb)))
        return st.value, nb.value

    def autotune(self, frames=2):
        st = ctypes.c_int()
        nb = ctypes.c_uint32()
        check(lib().lats_vbd_autotune(self._h, int(frames), ctypes.byref(st), ctypes.byref(nb)))
        return st.value, nb.value

    def autotune_results(self):
        """The last autotune's candidates: float array [n, 6] of (strategy,
        blocks or tiles, lanes, bands, best ms for the timed frames, colour
        block threads)."""
        n = ctypes.c_uint32()
        check(lib().lats_vbd_autotune_results(self._h, None, 0, ctypes.byref(n)))
        out = np.zeros((n.value, 6), np.float32)
        check(lib().lats_vbd_autotune_results(self._h, _p(out), n.value, ctypes.byref(n)))
        return out

    def set_obstacle(self, x, faces, thickness, stiffness, radius, cell):
        """Static triangle obstacle with penalty contact (see lats_vbd_set_obstacle)."""
        self._obs = (np.ascontiguousarray(x, np.float32), np.ascontiguousarray(faces, np.int32))
        x, f = self._obs
        check(lib().lats_vbd_set_obstacle(self._h, _p(x), len(x), _p(f), len(f), float(thickness),
                                          float(stiffness), float(radius), float(cell)))

    def frame_begin(self):
        check(lib().lats_vbd_frame_begin(self._h))

    def iterate(self, sweeps):
        check(lib().lats_vbd_iterate(self._h, int(sweeps)))

    def frame_end(self):
        check(lib().lats_vbd_frame_end(self._h))

    def energy_now(self):
        e = ctypes.c_float()
        check(lib().lats_vbd_energy_now(self._h, ctypes.byref(e)))
        return e.value

    def contacts(self):
        """(counts[V], face ids[V, 4] ascending, -1 padded) of the last detection."""
        n = self.cloth.num_verts
        c = np.zeros(n, np.uint32)
        f = np.zeros((n, 4), np.int32)
        check(lib().lats_vbd_get_contacts(self._h, _p(c), _p(f)))
        return c, f

    def set_energy_bending(self, on):
        check(lib().lats_vbd_set_energy_bending(self._h, int(bool(on))))

    def set_state(self, x, v):
        x = np.ascontiguousarray(x, np.float32)
        v = np.ascontiguousarray(v, np.float32)
        check(lib().lats_vbd_set_state(self._h, _p(x), _p(v)))

    def state(self):
        n = self.cloth.num_verts
        x = np.zeros((n, 3), np.float32)
        v = np.zeros((n, 3), np.float32)
        check(lib().lats_vbd_get_state(self._h, _p(x), _p(v)))
        return x, v

    def set_external_force(self, f):
        f = np.ascontiguousarray(f, np.float32)
        check(lib().lats_vbd_set_external_force(self._h, _p(f)))

    def physics_step(self, record_energy=False):
        check(lib().lats_vbd_physics_step(self._h, int(bool(record_energy))))
        return self.energy() if record_energy else None

    def energy(self):
        cnt = ctypes.c_uint32()
        check(lib().lats_vbd_get_energy(self._h, None, 0, ctypes.byref(cnt)))
        e = np.zeros(cnt.value, np.float32)
        if cnt.value:
            check(lib().lats_vbd_get_energy(self._h, _p(e), cnt.value, ctypes.byref(cnt)))
        return e

    def run_frames(self, num_frames, forces=None):
        fp = None
        if forces is not None:
            forces = np.ascontiguousarray(forces, np.float32).reshape(-1, 3)
            assert len(forces) >= num_frames
            fp = _p(forces)
        check(lib().lats_vbd_run_frames(self._h, int(num_frames), fp))

    def last_elapsed_ms(self):
        ms = ctypes.c_float()
        check(lib().lats_vbd_last_elapsed_ms(self._h, ctypes.byref(ms)))
        return ms.value

    def synchronize(self):
        check(lib().lats_vbd_synchronize(self._h))

    def dispatch(self, function_id, iter_idx=0, cluster_idx=0, buffer_idx=_lib.LATS_BUFFER_DEFAULT,
                 left_buffer_idx=_lib.LATS_BUFFER_NONE, right_buffer_idx=_lib.LATS_BUFFER_NONE,
                 input_buffer_idxs=(), is_allocated_to_main_device=True):
        p = _lib.LaunchParam()
        p.function_id = function_id
        p.iter_idx = iter_idx
        p.cluster_idx = cluster_idx
        p.buffer_idx = buffer_idx
        p.left_buffer_idx = left_buffer_idx
        p.right_buffer_idx = right_buffer_idx
        arr = (ctypes.c_uint32 * max(1, len(input_buffer_idxs)))(*input_buffer_idxs)
        p.num_input_buffers = len(input_buffer_idxs)
        p.input_buffer_idxs = arr
        p.is_allocated_to_main_device = int(bool(is_allocated_to_main_device))
        check(lib().lats_vbd_dispatch(self._h, ctypes.byref(p)))

    def set_record_energy(self, on):
        check(lib().lats_vbd_set_record_energy(self._h, int(bool(on))))

    def buffer(self, which, idx=0):
        """Copy of a position buffer (float32 [V, 3], vertex order); synchronises."""
        out = np.zeros((self.cloth.num_verts, 3), np.float32)
        check(lib().lats_vbd_buffer_download(self._h, int(which), int(idx), _p(out)))
        self.synchronize()
        return out

    def set_buffer(self, which, idx, values):
        """Overwrites a position buffer from float32 [V, 3] (vertex order); synchronises."""
        arr = np.ascontiguousarray(values, np.float32).reshape(self.cloth.num_verts, 3)
        check(lib().lats_vbd_buffer_upload(self._h, int(which), int(idx), _p(arr)))
        self.synchronize()

    def permutation(self):
        p = np.zeros(self.cloth.num_verts, np.uint32)
        check(lib().lats_vbd_copy_permutation(self._h, _p(p)))
        return p

    def time_colour_kernels(self, iterations):
        """(mean ms per colour-kernel launch, launches, vertex updates), HIP events."""
        ms = ctypes.c_float()
        n = ctypes.c_uint32()
        vu = ctypes.c_uint64()
        check(lib().lats_vbd_time_colour_kernels(self._h, int(iterations), ctypes.byref(ms), ctypes.byref(n),
                                                 ctypes.byref(vu)))
        return ms.value, n.value, vu.value

    def time_xpbd_kernels(self, iterations):
        """XPBD solver: (mean ms per spring-cluster launch, mean ms per
        bending-cluster launch, spring launches, bending launches)."""
        ms_s, ms_b = ctypes.c_float(), ctypes.c_float()
        n_s, n_b = ctypes.c_uint32(), ctypes.c_uint32()
        check(lib().lats_vbd_time_xpbd_kernels(self._h, int(iterations), ctypes.byref(ms_s), ctypes.byref(ms_b),
                                               ctypes.byref(n_s), ctypes.byref(n_b)))
        return ms_s.value, ms_b.value, n_s.value, n_b.value

    def device_bytes(self):
        b = ctypes.c_uint64()
        check(lib().lats_vbd_device_bytes(self._h, ctypes.byref(b)))
        return b.value


class SolverInterface:
    """Frame-level driver (example2_async_iter.cpp:1550-1710) over GpuVBD."""

    def __init__(self, cloth, params, **kw):
        self.cloth = cloth
        self.solver = GpuVBD(cloth, params, **kw)
        self.current_frame = 0
        self._saved = None

    def physics_step(self, record_energy=False):
        e = self.solver.physics_step(record_energy)
        self.current_frame += 1
        return e

    def restart_system(self):
        self.solver.set_state(self.cloth.x, np.zeros_like(self.cloth.x))
        self.current_frame = 0

    def save_current_frame_state(self):
        self._saved = self.solver.state()

    def load_saved_state(self):
        if self._saved is None:
            raise RuntimeError("no saved state")
        self.solver.set_state(*self._saved)

    def save_mesh_to_obj(self, path):
        """Writes the current positions as an OBJ (example2_async_iter.cpp:1649-1710)."""
        x, _ = self.solver.state()
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        with open(path, "w") as f:
            f.write("o mesh_0\n")
            for p in x:
                f.write(f"v {p[0]} {p[1]} {p[2]}\n")
            for t in self.cloth.faces + 1:
                f.write(f"f {t[0]} {t[1]} {t[2]}\n")


class GpuBatch:
    """Many independent cloths stepped together (one block per cloth)."""

    def __init__(self, cloths, params, stretch=None, bending=None):
        self.cloths = list(cloths)
        self.params = params
        n = len(self.cloths)
        descs = (_lib.MeshDesc * n)()
        self._keep = []
        for k, cl in enumerate(self.cloths):
            t = cl.topology
            self._keep += [cl.x, t.edges, t.bending, t.rest_length, t.bending_Q, t.adj_edges_csr,
                           t.adj_bending_csr, t.clusters_csr, cl.mass, cl.fixed]
            d = descs[k]
            d.num_verts, d.num_edges = t.num_verts, t.num_edges
            d.num_bending_edges, d.num_clusters = t.num_bending, t.num_clusters
            d.rest_x = _p(cl.x).value
            d.edges = _p(t.edges).value
            d.bending_edges = _p(t.bending).value
            d.edges_rest_length = _p(t.rest_length).value
            d.bending_Q = _p(t.bending_Q).value
            d.vert_adj_edges_csr = _p(t.adj_edges_csr).value
            d.vert_adj_bending_csr = _p(t.adj_bending_csr).value
            d.clusters_csr = _p(t.clusters_csr).value
            d.vert_mass = _p(cl.mass).value
            d.is_fixed = _p(cl.fixed).value
        ks = None if stretch is None else np.ascontiguousarray(stretch, np.float32)
        kb = None if bending is None else np.ascontiguousarray(bending, np.float32)
        self._params = params.c_struct()
        h = ctypes.c_void_p()
        check(lib().lats_batch_create(n, descs, None if ks is None else _p(ks), None if kb is None else _p(kb),
                                      ctypes.byref(self._params), ctypes.byref(h)))
        self._h = h
        _lib.track(self)
        del descs

    def close(self):
        if getattr(self, "_h", None):
            lib().lats_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def counts(self):
        n = ctypes.c_uint32()
        v = ctypes.c_uint64()
        check(lib().lats_batch_counts(self._h, ctypes.byref(n), ctypes.byref(v)))
        return n.value, v.value

    def state(self, k):
        nv = self.cloths[k].num_verts
        x = np.zeros((nv, 3), np.float32)
        v = np.zeros((nv, 3), np.float32)
        check(lib().lats_batch_get_state(self._h, k, _p(x), _p(v)))
        return x, v

    def set_state(self, k, x, v):
        x = np.ascontiguousarray(x, np.float32)
        v = np.ascontiguousarray(v, np.float32)
        check(lib().lats_batch_set_state(self._h, k, _p(x), _p(v)))

    def set_launch(self, block_threads, blocks=0):
        check(lib().lats_batch_set_launch(self._h, int(block_threads), int(blocks)))

    def launch(self):
        t = ctypes.c_uint32()
        b = ctypes.c_uint32()
        check(lib().lats_batch_get_launch(self._h, ctypes.byref(t), ctypes.byref(b)))
        return t.value, b.value

    def set_lanes_per_vertex(self, lanes):
        check(lib().lats_batch_set_lanes_per_vertex(self._h, int(lanes)))

    def lanes_per_vertex(self):
        n = ctypes.c_uint32()
        check(lib().lats_batch_get_lanes_per_vertex(self._h, ctypes.byref(n)))
        return n.value

    def set_lds_slots(self, slots):
        """Cloths of at most `slots` vertices keep their positions in LDS for
        the frame (0: none; at most lds_slots()[1])."""
        check(lib().lats_batch_set_lds_slots(self._h, int(slots)))

    def lds_slots(self):
        """(current LDS staging limit in vertices, the most this device allows)."""
        n = ctypes.c_uint32()
        mx = ctypes.c_uint32()
        check(lib().lats_batch_get_lds_slots(self._h, ctypes.byref(n), ctypes.byref(mx)))
        return n.value, mx.value

    def autotune(self, frames=1):
        t = ctypes.c_uint32()
        b = ctypes.c_uint32()
        check(lib().lats_batch_autotune(self._h, int(frames), ctypes.byref(t), ctypes.byref(b)))
        return t.value, b.value

    def run_frames(self, num_frames):
        check(lib().lats_batch_run_frames(self._h, int(num_frames)))

    def last_elapsed_ms(self):
        ms = ctypes.c_float()
        check(lib().lats_batch_last_elapsed_ms(self._h, ctypes.byref(ms)))
        return ms.value
