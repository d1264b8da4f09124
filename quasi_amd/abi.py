"""ctypes mirrors of the C ABI in include/quasi/*.h.

Field order and types follow the headers exactly (which in turn follow the
reference's src/quasi/plugin/plugin_interface.hpp:27-92 and
src/quasi/gpu/types.hpp:37-70).
"""
from __future__ import annotations

import ctypes as C

Q_PLUGIN_ABI_VERSION = 3

Q_GPU_BACKEND_NONE = 0
Q_GPU_BACKEND_METAL = 1
Q_GPU_BACKEND_VULKAN = 2
Q_GPU_BACKEND_WEBGPU = 3
Q_GPU_BACKEND_HIP = 4

Q_AOV_BEAUTY, Q_AOV_ALBEDO, Q_AOV_NORMAL, Q_AOV_DEPTH, Q_AOV_COUNT = 0, 1, 2, 3, 4

Q_EXT_OK = 0
Q_EXT_ERR_ARGUMENT = -1
Q_EXT_ERR_STATE = -2
Q_EXT_ERR_DEVICE = -3
Q_EXT_ERR_OUT_OF_MEMORY = -4
Q_EXT_ERR_UNSUPPORTED = -5
Q_NOALLOC_SCRATCH_BYTES = 512 << 20
Q_ADJOINT_PASS_PATHS = 1 << 24
Q_ADJOINT_NO_PARAM = 0xFFFFFFFF

Q_RNG_REFERENCE, Q_RNG_HASHED = 0, 1
Q_INTEGRATOR_REFERENCE, Q_INTEGRATOR_NEE = 0, 1
Q_ENGINE_MEGAKERNEL, Q_ENGINE_WAVEFRONT, Q_ENGINE_PERSISTENT = 0, 1, 2
Q_FLAG_RECORD_IDS = 1
Q_FLAG_TIME_KERNELS = 2
Q_STATS_MAX_KERNELS = 16

LOG_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_char_p)
SHUTDOWN_FN = C.CFUNCTYPE(None, C.c_void_p)


class Q_gpu_context(C.Structure):
    _fields_ = [("backend", C.c_uint32), ("device", C.c_void_p), ("queue", C.c_void_p), ("layer", C.c_void_p)]


class Q_camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("target", C.c_float * 3), ("up", C.c_float * 3), ("fov", C.c_float)]


class Q_render_frame(C.Structure):
    _fields_ = [("drawable", C.c_void_p), ("command_buffer", C.c_void_p), ("width", C.c_uint32),
                ("height", C.c_uint32), ("camera", Q_camera), ("camera_dirty", C.c_uint32)]


class Q_plugin_version(C.Structure):
    _fields_ = [("major", C.c_uint32), ("minor", C.c_uint32), ("patch", C.c_uint32)]


class Q_plugin_info(C.Structure):
    _fields_ = [("name", C.c_char_p), ("version", Q_plugin_version), ("description", C.c_char_p),
                ("author", C.c_char_p)]


class Q_plugin_context(C.Structure):
    _fields_ = [("viewport_width", C.c_uint32), ("viewport_height", C.c_uint32), ("host_data", C.c_void_p),
                ("gpu", C.POINTER(Q_gpu_context)), ("log", LOG_FN), ("request_shutdown", SHUTDOWN_FN)]


class Q_readback_result(C.Structure):
    _fields_ = [("data", C.POINTER(C.c_float)), ("width", C.c_uint32), ("height", C.c_uint32),
                ("channels", C.c_uint32)]


class Q_aov_buffer(C.Structure):
    _fields_ = [("data", C.POINTER(C.c_float)), ("width", C.c_uint32), ("height", C.c_uint32),
                ("channels", C.c_uint32)]


class Q_readback_aov_result(C.Structure):
    _fields_ = [("buffers", Q_aov_buffer * Q_AOV_COUNT)]


class Q_scene_desc(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("num_vertices", C.c_uint32), ("num_triangles", C.c_uint32),
                ("reserved", C.c_uint32), ("vertices", C.POINTER(C.c_float)), ("indices", C.POINTER(C.c_uint32)),
                ("albedo", C.POINTER(C.c_float)), ("emission", C.POINTER(C.c_float))]


class Q_render_params(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in (
        "struct_size", "spp_per_render", "max_depth", "rr_start", "seed", "rng_mode", "integrator", "engine",
        "bvh_width", "samples_per_pass", "block_size", "flags", "id_sample", "queue_chunk", "bvh_builder",
        "reserved")]


Q_BVH_BUILD_AUTO, Q_BVH_BUILD_HOST, Q_BVH_BUILD_GPU = 0, 1, 2
Q_QUEUE_FLAT = 0xFFFFFFFF   # queue_chunk: the plain (non-persistent) wavefront traversal
Q_QUEUE_LDS_TOP = 2         # queue_chunk: per-lane fetch, top BVH4 levels staged in LDS
Q_QUEUE_QUANT = 3           # queue_chunk: as Q_QUEUE_LDS_TOP over quantised 64-byte BVH4 nodes


class Q_kernel_stat(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("launches", C.c_uint64), ("total_ms", C.c_double)]


class Q_render_stats(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32),
                ("sample_count", C.c_uint32), ("num_triangles", C.c_uint32), ("num_lights", C.c_uint32),
                ("bvh_nodes", C.c_uint32), ("bvh_depth", C.c_uint32), ("rays_camera", C.c_uint64),
                ("rays_extension", C.c_uint64), ("rays_shadow", C.c_uint64), ("bvh_build_ms", C.c_double),
                ("render_ms", C.c_double), ("num_kernels", C.c_uint32), ("reserved", C.c_uint32),
                ("kernels", Q_kernel_stat * Q_STATS_MAX_KERNELS), ("bvh_builder", C.c_uint32),
                ("reserved2", C.c_uint32), ("bvh_sah_cost", C.c_double), ("bvh_device_ms", C.c_double),
                ("bvh_references", C.c_uint32), ("reserved3", C.c_uint32)]


class Q_batch_scene(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("seed", C.c_uint32), ("reserved", C.c_uint32),
                ("camera", Q_camera), ("mesh", Q_scene_desc)]


class Q_batch_desc(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("num_scenes", C.c_uint32), ("scenes", C.POINTER(Q_batch_scene))]


Q_AUTOTUNE_SINGLE, Q_AUTOTUNE_BATCH = 0, 1
Q_AUTOTUNE_USE_CACHE, Q_AUTOTUNE_NO_STORE = 1, 2
Q_AUTOTUNE_MAX_ENTRIES = 32


class Q_autotune_entry(C.Structure):
    _fields_ = [("bvh_width", C.c_uint32), ("block_size", C.c_uint32), ("queue_chunk", C.c_uint32),
                ("samples_per_pass", C.c_uint32), ("ms", C.c_double)]


class Q_autotune_result(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("num_entries", C.c_uint32), ("best", C.c_uint32),
                ("from_cache", C.c_uint32), ("entries", Q_autotune_entry * Q_AUTOTUNE_MAX_ENTRIES)]


Q_ADJOINT_IMAGE_GRADIENT, Q_ADJOINT_L2_TARGET = 0, 1
Q_ADJOINT_SEED = 1


class Q_adjoint_desc(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("mode", C.c_uint32), ("spp", C.c_uint32),
                ("sample_begin", C.c_uint32), ("seed", C.c_uint32), ("flags", C.c_uint32),
                ("image", C.POINTER(C.c_float)), ("grad_albedo", C.POINTER(C.c_double)),
                ("grad_vertices", C.POINTER(C.c_double)), ("max_pass_paths", C.c_uint32),
                ("num_albedo_params", C.c_uint32), ("albedo_param", C.POINTER(C.c_uint32))]


class Q_adjoint_result(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("reserved", C.c_uint32), ("paths", C.c_uint64),
                ("loss", C.c_double), ("kernel_ms", C.c_double), ("working_bytes", C.c_uint64),
                ("passes", C.c_uint32), ("bands", C.c_uint32)]


# Symbol names, plugin_interface.hpp:170-179
K_SYMBOL_ABI_VERSION = "Q_plugin_abi_version"
K_SYMBOL_GET_INFO = "Q_plugin_get_info"
K_SYMBOL_CREATE = "Q_plugin_create"
K_SYMBOL_DESTROY = "Q_plugin_destroy"
K_SYMBOL_UPDATE = "Q_plugin_update"
K_SYMBOL_RENDER = "Q_plugin_render"
K_SYMBOL_READBACK = "Q_plugin_readback"
K_SYMBOL_READBACK_FREE = "Q_plugin_readback_free"
K_SYMBOL_READBACK_AOV = "Q_plugin_readback_aov"
K_SYMBOL_READBACK_AOV_FREE = "Q_plugin_readback_aov_free"

REQUIRED_SYMBOLS = (K_SYMBOL_ABI_VERSION, K_SYMBOL_GET_INFO, K_SYMBOL_CREATE, K_SYMBOL_DESTROY,
                    K_SYMBOL_UPDATE, K_SYMBOL_RENDER)
OPTIONAL_SYMBOLS = (K_SYMBOL_READBACK, K_SYMBOL_READBACK_FREE, K_SYMBOL_READBACK_AOV,
                    K_SYMBOL_READBACK_AOV_FREE)
EXTENSION_SYMBOLS = ("Q_plugin_set_scene", "Q_plugin_get_params", "Q_plugin_set_params",
                     "Q_plugin_readback_ids", "Q_plugin_readback_display", "Q_plugin_get_stats", "Q_plugin_reset_stats",
                     "Q_plugin_synchronize", "Q_plugin_set_batch", "Q_plugin_render_batch",
                     "Q_plugin_readback_batch", "Q_plugin_readback_batch_ids", "Q_plugin_autotune",
                     "Q_plugin_adjoint")
