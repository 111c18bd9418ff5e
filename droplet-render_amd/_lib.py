"""ctypes binding of libdroplet_amd.so (the C ABI in include/droplet_render.h).

The shared library is built in-tree by build.py. Loading fails loudly when it
is missing: there is no CPU fallback for the render path.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# DRK_LIBRARY may name another build of the same library (the host-code
# sanitizer build of build.py, loaded by the CPU tests that check it).
LIB_PATH = os.environ.get("DRK_LIBRARY") or os.path.join(HERE, "libdroplet_amd.so")

DRK_OK = 0
DRK_EINVAL = -1
DRK_ENOMEM = -2
DRK_EDEVICE = -3
DRK_ELIMIT = -4

DRK_VOLUME_SDF = 0
DRK_VOLUME_FOG = 1
DRK_NO_BRICK = 0xFFFFFFFF

DRK_PHASE_MIE = 0
DRK_PHASE_HG = 1
DRK_ENV_HOSEK = 0
DRK_ENV_NULL = 1
DRK_ENV_MAP = 2
DRK_MAX_LIGHTS = 8
DRK_KERNEL_DEPTHCOMP = 0x1
DRK_WALK_AUTO = 0
DRK_WALK_OCTREE = 1
DRK_WALK_GRID = 2
DRK_ADJ_SEGMENT = 0
DRK_ADJ_NULL_COLLISION = 1


# Device-allocation gate: int gate(size_t bytes, void *user); non-zero refuses.
ALLOC_GATE = C.CFUNCTYPE(C.c_int, C.c_size_t, C.c_void_p)


class DrkError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("droplet-render error %d: %s" % (code, msg))
        self.code = code


class OctreeNode(C.Structure):
    """OctreeStructure (src/scene.h:21-31), 64 bytes."""
    _fields_ = [("ce", C.c_float * 4), ("chn", C.c_uint32 * 8), ("volx", C.c_uint32 * 2), ("qval", C.c_float * 2)]


class HostScene(C.Structure):
    _fields_ = [("ob", C.POINTER(OctreeNode)), ("node_count", C.c_uint32),
                ("pvol", C.POINTER(C.c_float) * 2), ("leafx", C.c_uint32 * 2),
                ("lvoxc", C.c_uint32), ("mlevel", C.c_uint32), ("root_ce", C.c_float * 4)]


class Sun(C.Structure):
    _fields_ = [("direction", C.c_float * 3), ("color", C.c_float * 3), ("angle", C.c_float)]


class SamplerDesc(C.Structure):
    _fields_ = [("phase", C.c_uint32), ("hg_g", C.c_float), ("env", C.c_uint32),
                ("sky_turbidity", C.c_float), ("sky_albedo", C.c_float),
                ("env_map", C.POINTER(C.c_float)), ("env_w", C.c_uint32), ("env_h", C.c_uint32),
                ("lights", C.POINTER(Sun)), ("light_count", C.c_uint32)]


assert C.sizeof(OctreeNode) == 64


class BatchJob(C.Structure):
    """drk_batch_job: one (kernel, tile) render of a batched launch."""
    _fields_ = [("kernel", C.c_void_p), ("x0", C.c_uint32), ("y0", C.c_uint32), ("tilex", C.c_uint32),
                ("tiley", C.c_uint32), ("samples", C.c_uint32), ("seed", C.c_uint64), ("sample_base", C.c_uint32),
                ("d_lights", C.c_void_p), ("d_env", C.c_void_p), ("d_sumsq_lights", C.c_void_p),
                ("d_sumsq_env", C.c_void_p)]


class ViewJob(C.Structure):
    """drk_view_job: one call of a views launch (record forward or adjoint)."""
    _fields_ = [("kernel", C.c_void_p), ("x0", C.c_uint32), ("y0", C.c_uint32), ("tilex", C.c_uint32),
                ("tiley", C.c_uint32), ("samples", C.c_uint32), ("seed", C.c_uint64), ("sample_base", C.c_uint32),
                ("d_lights", C.c_void_p), ("d_env", C.c_void_p), ("d_record", C.c_void_p),
                ("record_bytes", C.c_size_t)]

_lib = None

_P = C.c_void_p
_F = C.POINTER(C.c_float)
_U = C.POINTER(C.c_uint32)
_D = C.POINTER(C.c_double)
u32, u64, f32, f64 = C.c_uint32, C.c_uint64, C.c_float, C.c_double

_SIGS = {
    "drk_last_error": (C.c_char_p, []),
    "drk_device_info": (C.c_int, [C.c_char_p, u32, _U]),
    "drk_scene_build_dense": (C.c_int, [_F, _F, u32, u32, u32, f32, u32, C.POINTER(C.POINTER(HostScene))]),
    "drk_host_scene_free": (None, [C.POINTER(HostScene)]),
    "drk_scene_save": (C.c_int, [C.POINTER(HostScene), C.c_char_p]),
    "drk_scene_load": (C.c_int, [C.c_char_p, C.POINTER(C.POINTER(HostScene))]),
    "drk_scene_create": (C.c_int, [C.POINTER(OctreeNode), u32, _F, u32, _F, u32, u32, C.POINTER(_P)]),
    "drk_scene_update_fog": (C.c_int, [_P, _F, _F]),
    "drk_scene_update_fog_device": (C.c_int, [_P, _P, _P]),
    "drk_scene_destroy": (None, [_P]),
    "drk_kernel_create": (C.c_int, [_P, _F, _F, C.POINTER(SamplerDesc), _F, u32, f32, f32, u32, u32, u32, u32,
                                    u32, C.POINTER(_P)]),
    "drk_kernel_set_seed": (C.c_int, [_P, u64, u32]),
    "drk_kernel_set_tuning": (C.c_int, [_P, u32, u32]),
    "drk_kernel_set_timing": (C.c_int, [_P, C.c_int]),
    "drk_kernel_set_round_limit": (C.c_int, [_P, u32]),
    "drk_kernel_limit_hit": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "drk_batch_limit_hit": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "drk_set_alloc_gate": (C.c_int, [C.c_void_p, C.c_void_p]),
    "drk_kernel_timing": (C.c_int, [_P, C.POINTER(C.c_double), _U]),
    "drk_kernel_render": (C.c_int, [_P, u32, u32, u32, u32, u32]),
    "drk_kernel_shadow": (C.c_int, [_P, u32, u32, u32, u32, u32]),
    "drk_kernel_phb": (_F, [_P, u32]),
    "drk_kernel_destroy": (None, [_P]),
    "drk_kernel_render_device": (C.c_int, [_P, u32, u32, u32, u32, u32, u64, u32, _P, _P, _P, _P, _P]),
    "drk_kernel_render_device_noalloc": (C.c_int, [_P, u32, u32, u32, u32, u32, u64, u32, _P, _P, _P, _P, _P]),
    "drk_kernel_render_weighted_device": (C.c_int, [_P, u32, u32, u32, u32, u32, u64, u32, _P, _P, _P]),
    "drk_kernel_set_adjoint_estimator": (C.c_int, [_P, u32]),
    "drk_scene_set_reference_fog": (C.c_int, [_P, _P]),
    "drk_scene_set_walk": (C.c_int, [_P, u32]),
    "drk_scene_walk_info": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_float), C.POINTER(C.c_int)]),
    "drk_kernel_count": (C.c_int, [_P, u32, u32, u32, u32, u32, u64, u32, C.POINTER(C.c_ulonglong)]),
    "drk_kernel_adjoint": (C.c_int, [_P, u32, u32, u32, u32, u32, u64, u32, _F, _F, _F]),
    "drk_kernel_adjoint_device": (C.c_int, [_P, u32, u32, u32, u32, u32, u64, u32, _P, _P, _P, _P]),
    "drk_kernel_set_event_budget": (C.c_int, [_P, u64]),
    "drk_scene_events_begin": (C.c_int, [_P, u64, _P]),
    "drk_scene_events_end": (C.c_int, [_P, _P, _P]),
    "drk_kernel_footprint": (C.c_int, [_P, C.POINTER(C.c_size_t)]),
    "drk_scene_footprint": (C.c_int, [_P, C.POINTER(C.c_size_t)]),
    "drk_kernel_record_bytes": (C.c_int, [_P, u32, u32, u32, u32, u32, C.POINTER(C.c_size_t)]),
    "drk_kernel_render_record_device": (C.c_int, [_P, u32, u32, u32, u32, u32, u64, u32, _P, _P, _P, C.c_size_t,
                                                  _P]),
    "drk_kernel_adjoint_record_device": (C.c_int, [_P, u32, u32, u32, u32, u32, u64, u32, _P, _P, _P, _P,
                                                   C.c_size_t, _P]),
    "drk_batch_create": (C.c_int, [C.POINTER(BatchJob), u32, C.POINTER(_P)]),
    "drk_batch_set_tuning": (C.c_int, [_P, u32, f32]),
    "drk_batch_launch": (C.c_int, [_P, _P]),
    "drk_batch_launch_graph": (C.c_int, [_P, _P]),
    "drk_batch_autotune": (C.c_int, [_P, _P, u32, C.c_char_p, u32]),
    "drk_batch_destroy": (None, [_P]),
    "drk_views_render_record_device": (C.c_int, [C.POINTER(ViewJob), u32, _P]),
    "drk_views_adjoint_record_device": (C.c_int, [C.POINTER(ViewJob), u32, _P, _P]),
    "drk_synth_cloud": (C.c_int, [u32, u32, u64, _F, u32, _F]),
    "drk_probe_philox": (C.c_int, [_U, _U, u32, _U]),
    "drk_probe_ray_stream": (C.c_int, [_U, _U, _U, u32, u32, u32, _F]),
    "drk_probe_mie_eval": (C.c_int, [_F, u32, _F]),
    "drk_probe_mie_sample": (C.c_int, [_F, _F, _F, u32, _F]),
    "drk_probe_hg_eval": (C.c_int, [f32, _F, u32, _F]),
    "drk_probe_hg_sample": (C.c_int, [f32, _F, _F, _F, u32, _F]),
    "drk_probe_sun": (C.c_int, [C.POINTER(Sun), _F, _F, _F, u32, _F, _F, _F]),
    "drk_probe_sky": (C.c_int, [C.POINTER(Sun), f32, f32, _F, u32, _F]),
    "drk_probe_map_env": (C.c_int, [_F, u32, u32, _F, u32, _F]),
    "drk_probe_udiv": (C.c_int, [u32, u32, u32, C.c_uint64, C.POINTER(C.c_uint64)]),
    "drk_probe_hosek_state": (C.c_int, [f64, f64, f64, _D, _D]),
    "drk_probe_trilinear": (C.c_int, [_F, u32, _F, _F, u32, C.c_int, _F]),
    "drk_probe_log_unit": (C.c_int, [u32, _F, _F]),
    "drk_probe_traverse": (C.c_int, [_P, _F, _F, u32, u32, _U, _F, _F, _U]),
    "drk_probe_camera": (C.c_int, [_F, _F, u32, u32, _U, _U, u32, _F, _F, _U]),
}

# Every C-ABI entry point declared in include/*.h (checked by the CPU tests).
EXPORTED = sorted(k for k in _SIGS)


def lib():
    """Loads the in-tree library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libdroplet_amd.so is not built; run droplet-render_amd/build.py")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != DRK_OK:
        raise DrkError(rc, lib().drk_last_error().decode(errors="replace"))
    return rc


def fptr(a):
    """float* of a contiguous float32 numpy array (or None)."""
    if a is None:
        return None
    assert a.dtype.name == "float32" and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(_F)


def uptr(a):
    if a is None:
        return None
    assert a.dtype.name == "uint32" and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_U)


def dptr(a):
    assert a.dtype.name == "float64" and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_D)


_gate_keep = None


def set_alloc_gate(refuse):
    """Installs drk_set_alloc_gate: refuse(bytes) -> True makes that device
    allocation fail as out of memory. None removes the gate."""
    global _gate_keep
    if refuse is None:
        check(lib().drk_set_alloc_gate(None, None))
        _gate_keep = None
        return
    cb = ALLOC_GATE(lambda nbytes, user: 1 if refuse(int(nbytes)) else 0)
    check(lib().drk_set_alloc_gate(C.cast(cb, C.c_void_p), None))
    _gate_keep = cb
