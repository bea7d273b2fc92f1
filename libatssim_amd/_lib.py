"""ctypes binding of libatssim_hip.so (the C ABI in include/libatssim_hip.h).

The library is built in-tree (`make -C libatssim_amd/csrc`, or
`__graft_entry__.build()`).  Loading fails loudly when it is missing: there
is no CPU fallback for the product path.
"""
import atexit
import ctypes
import os
import weakref

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libatssim_hip.so")

LATS_OK = 0
LATS_ERR_INVALID_ARGUMENT = 1
LATS_ERR_HIP = 2
LATS_ERR_NO_DEVICE = 3
LATS_ERR_UNSUPPORTED = 4
LATS_ERR_OUT_OF_MEMORY = 5

LATS_TOPO_EDGES = 0
LATS_TOPO_BENDING = 1
LATS_TOPO_REST_LENGTH = 2
LATS_TOPO_BENDING_Q = 3
LATS_TOPO_ADJ_EDGES_CSR = 4
LATS_TOPO_ADJ_BENDING_CSR = 5
LATS_TOPO_CLUSTERS_CSR = 6
LATS_TOPO_COLOUR = 7
LATS_TOPO_FACES = 8

LATS_STRATEGY_AUTO = 0
LATS_STRATEGY_GRAPH = 1
LATS_STRATEGY_PERSISTENT = 2
LATS_STRATEGY_FLOW = 3
LATS_STRATEGY_PATCH = 4

LATS_FN_ADDITIONAL_ROOT = 0
LATS_FN_ADDITIONAL_TERMINAL = 1
LATS_FN_PREDICT_POSITION = 2
LATS_FN_VBD_ALL_IN_ONE = 3
LATS_FN_CONSTRAINT_LAST_NODE = 4
LATS_FN_UPDATE_VELOCITY = 5
LATS_FN_RESET_CONSTRAINS = 6
LATS_FN_XPBD_STRETCH = 7
LATS_FN_XPBD_BENDING = 8
LATS_SOLVER_VBD, LATS_SOLVER_XPBD = 0, 1

LATS_BUFFER_DEFAULT = 0x0FFFFFFF
LATS_BUFFER_NONE = 0xFFFFFFFF


class LatsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libatssim_hip error {code}: {msg}")
        self.code = code


class SceneParams(ctypes.Structure):
    _fields_ = [
        ("implicit_dt", ctypes.c_float),
        ("num_substep", ctypes.c_uint32),
        ("constraint_iter_count", ctypes.c_uint32),
        ("stiffness_stretch_spring", ctypes.c_float),
        ("stiffness_quadratic_bending", ctypes.c_float),
        ("damping_cloth", ctypes.c_float),
        ("gravity", ctypes.c_float * 3),
    ]


class MeshDesc(ctypes.Structure):
    _fields_ = [
        ("num_verts", ctypes.c_uint32),
        ("num_edges", ctypes.c_uint32),
        ("num_bending_edges", ctypes.c_uint32),
        ("num_clusters", ctypes.c_uint32),
        ("rest_x", ctypes.c_void_p),
        ("edges", ctypes.c_void_p),
        ("bending_edges", ctypes.c_void_p),
        ("edges_rest_length", ctypes.c_void_p),
        ("bending_Q", ctypes.c_void_p),
        ("vert_adj_edges_csr", ctypes.c_void_p),
        ("vert_adj_bending_csr", ctypes.c_void_p),
        ("clusters_csr", ctypes.c_void_p),
        ("vert_mass", ctypes.c_void_p),
        ("is_fixed", ctypes.c_void_p),
    ]


# position buffers across processors (lats_vbd_buffer_download / _upload)
LATS_BUF_ITER, LATS_BUF_BEGIN, LATS_BUF_X, LATS_BUF_X_START, LATS_BUF_V = range(5)


class LaunchParam(ctypes.Structure):
    _fields_ = [
        ("function_id", ctypes.c_uint32),
        ("iter_idx", ctypes.c_uint32),
        ("cluster_idx", ctypes.c_uint32),
        ("buffer_idx", ctypes.c_uint32),
        ("left_buffer_idx", ctypes.c_uint32),
        ("right_buffer_idx", ctypes.c_uint32),
        ("num_input_buffers", ctypes.c_uint32),
        ("input_buffer_idxs", ctypes.POINTER(ctypes.c_uint32)),
        ("is_allocated_to_main_device", ctypes.c_uint32),
    ]


class PatchInfo(ctypes.Structure):
    _fields_ = [
        ("patches", ctypes.c_uint32),
        ("colours_per_launch", ctypes.c_uint32),
        ("launches_per_sweep", ctypes.c_uint32),
        ("threads", ctypes.c_uint32),
        ("lanes", ctypes.c_uint32),
        ("max_region", ctypes.c_uint32),
        ("updates_per_sweep", ctypes.c_uint64),
        ("region_slots_per_sweep", ctypes.c_uint64),
        ("ell_entries_per_sweep", ctypes.c_uint64),
        ("build_ms", ctypes.c_double),
    ]


# name -> (restype, argtypes); every symbol declared in include/libatssim_hip.h
_VP = ctypes.c_void_p
_U32 = ctypes.c_uint32
_I = ctypes.c_int
SIGNATURES = {
    "lats_last_error": (ctypes.c_char_p, []),
    "lats_topology_build": (_I, [_VP, _U32, _VP, _U32, ctypes.POINTER(_VP)]),
    "lats_obj_read": (_I, [ctypes.c_char_p, ctypes.POINTER(_VP)]),
    "lats_obj_parse": (_I, [ctypes.c_char_p, ctypes.POINTER(_VP)]),
    "lats_obj_counts": (_I, [_VP, _VP]),
    "lats_obj_copy": (_I, [_VP, _VP, _VP]),
    "lats_obj_destroy": (None, [_VP]),
    "lats_obj_write": (_I, [ctypes.c_char_p, _VP, _U32, _VP, _U32, _U32]),
    "lats_topology_destroy": (None, [_VP]),
    "lats_topology_save": (_I, [_VP, ctypes.c_char_p]),
    "lats_topology_load": (_I, [ctypes.c_char_p, ctypes.POINTER(_VP)]),
    "lats_topology_counts": (_I, [_VP, _VP]),
    "lats_topology_size": (_I, [_VP, _I, ctypes.POINTER(ctypes.c_uint64)]),
    "lats_topology_copy": (_I, [_VP, _I, _VP]),
    "lats_scene_grid": (_I, [_U32, _U32, ctypes.c_float, ctypes.c_float, ctypes.c_uint64, _VP, _VP]),
    "lats_scene_wind": (_I, [_U32, _U32, ctypes.c_float, ctypes.c_uint64, _VP]),
    "lats_vbd_create": (_I, [ctypes.POINTER(MeshDesc), ctypes.POINTER(SceneParams), ctypes.POINTER(_VP)]),
    "lats_vbd_destroy": (_I, [_VP]),
    "lats_vbd_set_strategy": (_I, [_VP, _I]),
    "lats_vbd_set_energy_bending": (_I, [_VP, _I]),
    "lats_vbd_set_persistent_blocks": (_I, [_VP, _U32]),
    "lats_vbd_set_lanes_per_vertex": (_I, [_VP, _U32]),
    "lats_vbd_set_regions": (_I, [_VP, _U32]),
    "lats_vbd_set_patches": (_I, [_VP, _U32, _U32, _U32]),
    "lats_vbd_get_patches": (_I, [_VP, ctypes.POINTER(PatchInfo)]),
    "lats_vbd_set_solver": (_I, [_VP, _I]),
    "lats_vbd_get_solver": (_I, [_VP, ctypes.POINTER(_I)]),
    "lats_vbd_xpbd_clusters": (_I, [_VP, _I, _VP, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]),
    "lats_vbd_get_regions": (_I, [_VP, ctypes.POINTER(_U32)]),
    "lats_vbd_set_colour_block": (_I, [_VP, _U32]),
    "lats_vbd_get_colour_block": (_I, [_VP, ctypes.POINTER(_U32)]),
    "lats_vbd_frame_begin": (_I, [_VP]),
    "lats_vbd_iterate": (_I, [_VP, _U32]),
    "lats_vbd_frame_end": (_I, [_VP]),
    "lats_vbd_energy_now": (_I, [_VP, ctypes.POINTER(ctypes.c_float)]),
    "lats_vbd_energy_noalloc": (_I, [_VP, _VP, _VP]),
    "lats_vbd_set_obstacle": (_I, [_VP, _VP, _U32, _VP, _U32, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                   ctypes.c_float]),
    "lats_vbd_get_contacts": (_I, [_VP, _VP, _VP]),
    "lats_scene_icosphere": (_I, [_U32, ctypes.c_float, _VP, _VP, _VP, ctypes.POINTER(_U32), ctypes.POINTER(_U32)]),
    "lats_vbd_get_lanes_per_vertex": (_I, [_VP, ctypes.POINTER(_U32)]),
    "lats_vbd_get_strategy": (_I, [_VP, ctypes.POINTER(_I), ctypes.POINTER(_U32)]),
    "lats_vbd_autotune": (_I, [_VP, _U32, ctypes.POINTER(_I), ctypes.POINTER(_U32)]),
    "lats_vbd_set_state": (_I, [_VP, _VP, _VP]),
    "lats_vbd_get_state": (_I, [_VP, _VP, _VP]),
    "lats_vbd_set_external_force": (_I, [_VP, _VP]),
    "lats_vbd_physics_step": (_I, [_VP, _I]),
    "lats_vbd_get_energy": (_I, [_VP, _VP, _U32, ctypes.POINTER(_U32)]),
    "lats_vbd_run_frames": (_I, [_VP, _U32, _VP]),
    "lats_vbd_last_elapsed_ms": (_I, [_VP, ctypes.POINTER(ctypes.c_float)]),
    "lats_vbd_synchronize": (_I, [_VP]),
    "lats_vbd_dispatch": (_I, [_VP, ctypes.POINTER(LaunchParam)]),
    "lats_vbd_set_record_energy": (_I, [_VP, _I]),
    "lats_vbd_stream_wait_host": (_I, [_VP, _U32]),
    "lats_vbd_host_signal": (_I, [_VP, _U32]),
    "lats_vbd_record_marker": (_I, [_VP, _U32]),
    "lats_vbd_wait_marker": (_I, [_VP, _U32]),
    "lats_vbd_marker_elapsed_ms": (_I, [_VP, _U32, _U32, ctypes.POINTER(ctypes.c_float)]),
    "lats_vbd_query_marker": (_I, [_VP, _U32, ctypes.POINTER(_I)]),
    "lats_vbd_buffer_download": (_I, [_VP, _I, _U32, _VP]),
    "lats_vbd_buffer_upload": (_I, [_VP, _I, _U32, _VP]),
    "lats_host_alloc": (_I, [ctypes.c_size_t, ctypes.POINTER(_VP)]),
    "lats_host_free": (_I, [_VP]),
    "lats_vbd_device_positions": (_I, [_VP, ctypes.POINTER(_VP)]),
    "lats_vbd_copy_permutation": (_I, [_VP, _VP]),
    "lats_vbd_get_stream": (_I, [_VP, ctypes.POINTER(_VP)]),
    "lats_vbd_time_colour_kernels": (_I, [_VP, _U32, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(_U32),
                                          ctypes.POINTER(ctypes.c_uint64)]),
    "lats_vbd_device_bytes": (_I, [_VP, ctypes.POINTER(ctypes.c_uint64)]),
    "lats_batch_create": (_I, [_U32, _VP, _VP, _VP, ctypes.POINTER(SceneParams), ctypes.POINTER(_VP)]),
    "lats_batch_destroy": (_I, [_VP]),
    "lats_batch_counts": (_I, [_VP, ctypes.POINTER(_U32), ctypes.POINTER(ctypes.c_uint64)]),
    "lats_batch_set_state": (_I, [_VP, _U32, _VP, _VP]),
    "lats_batch_get_state": (_I, [_VP, _U32, _VP, _VP]),
    "lats_batch_set_launch": (_I, [_VP, _U32, _U32]),
    "lats_batch_get_launch": (_I, [_VP, ctypes.POINTER(_U32), ctypes.POINTER(_U32)]),
    "lats_batch_autotune": (_I, [_VP, _U32, ctypes.POINTER(_U32), ctypes.POINTER(_U32)]),
    "lats_batch_run_frames": (_I, [_VP, _U32]),
    "lats_batch_last_elapsed_ms": (_I, [_VP, ctypes.POINTER(ctypes.c_float)]),
    "lats_batch_set_lanes_per_vertex": (_I, [_VP, _U32]),
    "lats_batch_get_lanes_per_vertex": (_I, [_VP, ctypes.POINTER(_U32)]),
    "lats_batch_set_lds_slots": (_I, [_VP, _U32]),
    "lats_batch_get_lds_slots": (_I, [_VP, ctypes.POINTER(_U32), ctypes.POINTER(_U32)]),
    "lats_mlp_create": (_I, [ctypes.c_uint64, ctypes.POINTER(_VP)]),
    "lats_mlp_destroy": (_I, [_VP]),
    "lats_mlp_num_params": (_I, [ctypes.POINTER(_U32)]),
    "lats_mlp_get_params": (_I, [_VP, _VP]),
    "lats_mlp_set_params": (_I, [_VP, _VP]),
    "lats_mlp_get_feature_scale": (_I, [_VP, _VP]),
    "lats_mlp_set_feature_scale": (_I, [_VP, _VP]),
    "lats_mlp_forward_host": (_I, [_VP, _VP, _U32, _VP]),
    "lats_mlp_gradient_host": (_I, [_VP, _VP, _VP, _U32, _VP, ctypes.POINTER(ctypes.c_double)]),
    "lats_mlp_bind": (_I, [_VP, _VP]),
    "lats_mlp_reset_history": (_I, [_VP]),
    "lats_mlp_set_stop_rule": (_I, [_VP, ctypes.c_float]),
    "lats_mlp_set_guess_options": (_I, [_VP, _U32, _I]),
    "lats_mlp_reserve_rows": (_I, [_VP, ctypes.c_uint64]),
    "lats_mlp_record_frame": (_I, [_VP, _VP]),
    "lats_mlp_dataset_append": (_I, [_VP, _VP, _VP, _U32]),
    "lats_mlp_dataset_rows": (_I, [_VP, ctypes.POINTER(ctypes.c_uint64)]),
    "lats_mlp_dataset_read": (_I, [_VP, ctypes.c_uint64, _U32, _VP, _VP]),
    "lats_mlp_dataset_clear": (_I, [_VP]),
    "lats_mlp_fit_feature_scale": (_I, [_VP, _VP]),
    "lats_mlp_time_kernels": (_I, [_VP, _U32, _U32, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
    "lats_mlp_train": (_I, [_VP, _U32, _U32, ctypes.c_float, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]),
    "lats_mlp_warm_start_frame": (_I, [_VP, _VP, ctypes.c_float, _U32, _U32, _U32, _I, _VP]),
    "lats_mlp_record_guided_frame": (_I, [_VP, _VP, _VP, _U32, _U32, _U32, _I, _VP]),
    "lats_vbd_time_xpbd_kernels": (_I, [_VP, _U32, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                        ctypes.POINTER(_U32), ctypes.POINTER(_U32)]),
    "lats_vbd_autotune_results": (_I, [_VP, _VP, _U32, ctypes.POINTER(_U32)]),
}

_LIB = None


def lib():
    """The loaded C-ABI library (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `make -C libatssim_amd/csrc` "
                "(the HIP path has no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


# Every live context, batch and network, closed by an atexit hook: Python runs
# it before the C runtime's exit handlers (and so before the HIP runtime tears
# itself down), whereas objects left to __del__ at interpreter teardown may be
# destroyed after that.  Networks close first (they hold bound contexts).
_LIVE = weakref.WeakSet()


def track(obj):
    """Registers an object with a close() method for closing at exit."""
    _LIVE.add(obj)


def close_all():
    """Closes every tracked object that is still open."""
    objs = sorted(list(_LIVE), key=lambda o: getattr(o, "_close_rank", 1))
    for o in objs:
        o.close()


atexit.register(close_all)


MEASURE_PATH = os.path.join(_HERE, "lib", "liblats_measure.so")
LATS_FLOOR_EMPTY, LATS_FLOOR_GATHER, LATS_FLOOR_GRID_BARRIER, LATS_FLOOR_REREAD = 0, 1, 2, 3
LATS_FLOOR_HANDOFF_FENCED, LATS_FLOOR_HANDOFF_RELAXED, LATS_FLOOR_HANDOFF_XCD = 4, 5, 6
_MEASURE = None


def measure_lib():
    """The benchmark's latency-floor library (liblats_measure.so, built beside
    the solver library); lats_time_floor reports errors via lats_last_error."""
    global _MEASURE
    if _MEASURE is None:
        lib()
        if not os.path.exists(MEASURE_PATH):
            raise RuntimeError(f"{MEASURE_PATH} is missing: build it with `make -C libatssim_amd/csrc`")
        M = ctypes.CDLL(MEASURE_PATH)
        M.lats_time_floor.restype = _I
        M.lats_time_floor.argtypes = [_I, _U32, _U32, ctypes.POINTER(ctypes.c_float)]
        _MEASURE = M
    return _MEASURE


def check(rc):
    if rc != LATS_OK:
        msg = lib().lats_last_error()
        raise LatsError(rc, msg.decode() if msg else "")
    return rc
