"""ctypes binding of libmeso_hip.so (the C-ABI declared in include/meso_hip.h).

The library is loaded from mesoengine_amd/_build/.  There is no CPU fallback:
if the library is missing or no GPU is present, the calls fail loudly.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libmeso_hip.so")

OK, ERR_ARG, ERR_RUNTIME, ERR_OOM, ERR_BUSY, ERR_NOT_SUPPORTED = range(6)
SCENE_DENSE_U8, SCENE_BRICKMAP_U16, SCENE_TWO_LEVEL_U16 = 1, 2, 3
GEN_SPHERES, GEN_HEIGHTFIELD, GEN_TERRAIN_CAVES, GEN_CAVES = 0, 1, 2, 3
PART_DENSE, PART_SLOTS, PART_MASKS, PART_VOXELS, PART_TOP = range(5)
RAY_PRIMARY, RAY_SHADOW, RAY_AO, RAY_COUNT = 0, 1, 2, 0x100
MESH_CULLED, MESH_GREEDY = 0, 1
CHUNKS_PACKED, CHUNKS_DENSE_GRID = 0, 1
CNT_NAMES = ["n_vox", "n_top", "n_node", "n_brickslot", "n_brickmask", "hits", "rays",
             "shadow_rays", "ao_rays", "n_coarse", "n_iter"]


class MesoError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("meso status %d: %s" % (status, msg))
        self.status = status


class DeviceInfo(C.Structure):
    _fields_ = [("size", C.c_uint32), ("name", C.c_char * 64), ("compute_units", C.c_int32),
                ("xcds", C.c_int32), ("lds_bytes_per_cu", C.c_int32), ("l2_bytes", C.c_int32),
                ("hbm_bytes", C.c_uint64), ("clock_khz", C.c_int32), ("wavefront", C.c_int32)]


class SceneInfo(C.Structure):
    _fields_ = [("size", C.c_uint32), ("kind", C.c_int32), ("n", C.c_int32),
                ("n_bricks", C.c_int64), ("n_nodes", C.c_int64), ("device_bytes", C.c_uint64),
                ("solid_voxels", C.c_int64)]


class GenerateDesc(C.Structure):
    _fields_ = [("size", C.c_uint32), ("generator", C.c_int32), ("layout", C.c_int32),
                ("n", C.c_int32), ("seed", C.c_uint64), ("n_spheres", C.c_int32)]


class Camera(C.Structure):
    _fields_ = [("size", C.c_uint32), ("eye", C.c_float * 3), ("target", C.c_float * 3),
                ("up", C.c_float * 3), ("vfov_deg", C.c_float), ("width", C.c_int32),
                ("height", C.c_int32)]


class Shading(C.Structure):
    _fields_ = [("size", C.c_uint32), ("light", C.c_float * 3), ("ao_seed", C.c_uint32),
                ("ao_samples", C.c_int32), ("ao_radius", C.c_float)]


class RayOutputs(C.Structure):
    _fields_ = [("size", C.c_uint32), ("hit_index", C.c_void_p), ("index_bytes", C.c_int32),
                ("face", C.c_void_p), ("t", C.c_void_p), ("shade", C.c_void_p),
                ("counters", C.c_void_p)]


class ChunkSource(C.Structure):
    _fields_ = [("size", C.c_uint32), ("voxels", C.c_void_p), ("elem_bytes", C.c_int32),
                ("layout", C.c_int32), ("grid_n", C.c_int32), ("n_chunks", C.c_int32)]


class TuneResult(C.Structure):
    _fields_ = [("size", C.c_uint32), ("variant", C.c_int32), ("block", C.c_int32),
                ("grid", C.c_int32), ("tile_w", C.c_int32), ("tile_h", C.c_int32),
                ("best_ms", C.c_float), ("default_ms", C.c_float), ("tried", C.c_int32)]


class ServeConfig(C.Structure):
    _fields_ = [("size", C.c_uint32), ("hbm_budget_bytes", C.c_uint64),
                ("max_pending", C.c_int32), ("max_inflight", C.c_int32),
                ("workgroups_per_cu", C.c_int32), ("tile", C.c_int32),
                ("idle_exit_us", C.c_int32)]


class ServeStats(C.Structure):
    _fields_ = [("size", C.c_uint32), ("submitted", C.c_uint64), ("admitted", C.c_uint64),
                ("completed", C.c_uint64), ("refused_busy", C.c_uint64),
                ("launches", C.c_uint64), ("primary_rays", C.c_uint64),
                ("bytes_in_use", C.c_uint64), ("peak_bytes_in_use", C.c_uint64),
                ("busy_ms", C.c_double), ("queue_us_sum", C.c_double),
                ("queue_us_max", C.c_double), ("queued_over_1ms", C.c_uint64),
                ("admission_blocked_passes", C.c_uint64)]


class HostOutputs(C.Structure):
    _fields_ = [("size", C.c_uint32), ("index_bytes", C.c_int32), ("hit_index", C.c_void_p),
                ("face", C.c_void_p), ("t", C.c_void_p), ("shade", C.c_void_p)]


class PoolDesc(C.Structure):
    _fields_ = [("size", C.c_uint32), ("capacity", C.c_int32), ("quads_per_slot", C.c_int32),
                ("forward_radius", C.c_int32), ("backward_radius", C.c_int32),
                ("view_angle", C.c_float), ("max_loads", C.c_int32), ("mesh_mode", C.c_uint32),
                ("bake_directions", C.c_int32)]


class PoolStats(C.Structure):
    _fields_ = [("size", C.c_uint32), ("stamp", C.c_uint32), ("n_candidates", C.c_uint32),
                ("n_loads", C.c_uint32), ("n_evicted", C.c_uint32), ("n_resident", C.c_uint32),
                ("n_truncated", C.c_uint32), ("pad", C.c_uint32), ("n_quads", C.c_uint64),
                ("device_bytes", C.c_uint64)]


class PoolArrays(C.Structure):
    _fields_ = [("size", C.c_uint32), ("chunks_per_axis", C.c_int32), ("lookup", C.c_void_p),
                ("slot_chunk", C.c_void_p), ("slot_stamp", C.c_void_p),
                ("slot_count", C.c_void_p), ("quads", C.c_void_p), ("load_xyz", C.c_void_p),
                ("load_slot", C.c_void_p)]


class RasterStats(C.Structure):
    _fields_ = [("size", C.c_uint32), ("small_quads", C.c_uint32), ("large_quads", C.c_uint32),
                ("segments", C.c_uint32), ("overflow_quads", C.c_uint32)]


P = C.c_void_p
PP = C.POINTER(C.c_void_p)
_SIGS = {
    "meso_last_error_message": (C.c_char_p, []),
    "meso_abi_version": (C.c_uint32, []),
    "meso_context_create": (C.c_int32, [C.c_int32, PP]),
    "meso_context_destroy": (C.c_int32, [P]),
    "meso_device_info_get": (C.c_int32, [P, P]),
    "meso_buffer_create": (C.c_int32, [P, C.c_uint64, P, PP]),
    "meso_buffer_destroy": (C.c_int32, [P]),
    "meso_buffer_upload": (C.c_int32, [P, P, C.c_uint64, C.c_uint64]),
    "meso_buffer_download": (C.c_int32, [P, P, C.c_uint64, C.c_uint64]),
    "meso_buffer_device_address": (C.c_uint64, [P, C.c_uint64]),
    "meso_buffer_size": (C.c_uint64, [P]),
    "meso_buffer_fill_zero": (C.c_int32, [P, P]),
    "meso_stream_create": (C.c_int32, [P, PP]),
    "meso_stream_destroy": (C.c_int32, [P, P]),
    "meso_stream_sync": (C.c_int32, [P, P]),
    "meso_event_create": (C.c_int32, [P, PP]),
    "meso_event_destroy": (C.c_int32, [P]),
    "meso_event_record": (C.c_int32, [P, P]),
    "meso_stream_wait_event": (C.c_int32, [P, P, P]),
    "meso_event_sync": (C.c_int32, [P]),
    "meso_event_elapsed_ms": (C.c_int32, [P, P, C.POINTER(C.c_float)]),
    "meso_memcpy_d2h": (C.c_int32, [P, P, P, C.c_uint64, P]),
    "meso_memcpy_h2d": (C.c_int32, [P, P, P, C.c_uint64, P]),
    "meso_memcpy_d2d": (C.c_int32, [P, P, P, C.c_uint64, P]),
    "meso_memcpy_h2d_async": (C.c_int32, [P, P, P, C.c_uint64, P]),
    "meso_host_alloc": (C.c_int32, [P, C.c_uint64, P]),
    "meso_host_free": (C.c_int32, [P, P]),
    "meso_scene_create_dense_u8": (C.c_int32, [P, C.c_int32, P, PP]),
    "meso_scene_create_brickmap_u16": (C.c_int32, [P, C.c_int32, P, C.c_int64, P, P, PP]),
    "meso_scene_create_two_level_u16": (C.c_int32, [P, C.c_int32, P, C.c_int64, P, C.c_int64,
                                                    P, P, PP]),
    "meso_scene_generate": (C.c_int32, [P, P, PP]),
    "meso_scene_get_info": (C.c_int32, [P, P]),
    "meso_scene_download": (C.c_int32, [P, C.c_int32, P, C.c_uint64]),
    "meso_scene_update_bricks": (C.c_int32, [P, P, P, C.c_int32, P]),
    "meso_scene_update_bricks_async": (C.c_int32, [P, P, P, C.c_int32, P, P]),
    "meso_scene_edit_chunks": (C.c_int32, [P, P, C.c_int32, C.c_uint32, P]),
    "meso_mesh_scene_chunks": (C.c_int32, [P, P, P, C.c_int32, C.c_uint32, P, P, P, C.c_uint64,
                                           P, P, P]),
    "meso_scene_destroy": (C.c_int32, [P]),
    "meso_raymarch": (C.c_int32, [P, P, P, C.c_uint32, P, P, P]),
    "meso_mesh_chunk": (C.c_int32, [P, P, C.c_int32, C.c_uint32, P, P, C.c_uint32, P]),
    "meso_mesh_workspace_bytes": (C.c_uint64, [C.c_int32]),
    "meso_mesh_chunks": (C.c_int32, [P, P, C.c_uint32, P, P, P, C.c_uint64, P, P, P]),
    "meso_mesh_chunks_noalloc": (C.c_int32, [P, P, C.c_uint32, P, P, P, C.c_uint64, P, P]),
    "meso_mesh_check": (C.c_int32, [P, P, P]),
    "meso_mesh_set_lookback_spin_limit": (C.c_int32, [P, C.c_uint32]),
    "meso_mesh_graph_create": (C.c_int32, [P, P, C.c_uint32, P, P, P, C.c_uint64, P, P, PP]),
    "meso_mesh_plain_graph_create": (C.c_int32, [P, P, C.c_uint32, P, P, C.c_uint32, PP]),
    "meso_graph_launch": (C.c_int32, [P, P]),
    "meso_graph_destroy": (C.c_int32, [P]),
    "meso_generate_chunks": (C.c_int32, [P, C.c_uint64, C.c_int32, P, P]),
    "meso_generate_sphere_chunks": (C.c_int32, [P, P, C.c_int32, C.c_int32, C.c_float,
                                                C.c_int32, P, P]),
    "meso_autotune_mesh": (C.c_int32, [P, P, C.c_uint32, P, P, P, C.c_uint64, P, P, P]),
    "meso_autotune_raymarch": (C.c_int32, [P, P, P, C.c_uint32, P, P, P]),
    "meso_autotune_pin_raymarch": (C.c_int32, [P, P, P, C.c_uint32, P]),
    "meso_autotune_load": (C.c_int32, [P, C.c_char_p]),
    "meso_autotune_save": (C.c_int32, [P, C.c_char_p]),
    "meso_serve_start": (C.c_int32, [P, P, PP]),
    "meso_serve_submit": (C.c_int32, [P, P, P, C.c_uint32, P, C.POINTER(C.c_uint64)]),
    "meso_serve_poll": (C.c_int32, [P, C.c_uint64, C.POINTER(C.c_int32)]),
    "meso_serve_wait": (C.c_int32, [P, C.c_uint64, C.c_int64, C.POINTER(C.c_double)]),
    "meso_serve_take": (C.c_int32, [P, C.c_uint64, P]),
    "meso_serve_stats_get": (C.c_int32, [P, P]),
    "meso_serve_stop": (C.c_int32, [P]),
    "meso_chunk_pool_create": (C.c_int32, [P, P, P, PP]),
    "meso_chunk_pool_destroy": (C.c_int32, [P]),
    "meso_chunk_pool_update": (C.c_int32, [P, P, P, P]),
    "meso_chunk_pool_remesh": (C.c_int32, [P, P, C.c_int32, P]),
    "meso_chunk_pool_stats": (C.c_int32, [P, P, P]),
    "meso_chunk_pool_arrays": (C.c_int32, [P, P]),
    "meso_raster_create": (C.c_int32, [P, C.c_int32, C.c_int32, PP]),
    "meso_raster_destroy": (C.c_int32, [P]),
    "meso_raster_pool": (C.c_int32, [P, P, P, P, P, P]),
    "meso_raster_stats_get": (C.c_int32, [P, P, P]),
}
EXPORTED = sorted(_SIGS)

_lib = None


def load():
    """Load libmeso_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libmeso_hip.so not built (%s); run __graft_entry__.build()"
                               % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != OK:
        msg = load().meso_last_error_message()
        raise MesoError(status, msg.decode() if msg else "")
    return status


def byref(s):
    """Pointer argument to a struct for a void-pointer parameter.  The
    returned object holds a reference to `s`, so a temporary passed as
    byref(make_struct()) lives until the call returns.  (A c_void_p made by
    casting holds none: the struct could be freed, and its first bytes
    reused by the allocator, before the library read it.)"""
    return C.byref(s)
