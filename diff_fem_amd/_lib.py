"""ctypes binding of libpdm.so (the C ABI of include/pdm.h).

The library is built in-tree (``diff_fem_amd/libpdm.so``) by
``diff_fem_amd.build.build()``. There is no fallback: if the shared library
is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpdm.so")

PDM_FP64 = 0
PDM_MIXED = 1
PDM_MIXED_ADJOINT = 2
PDM_HOST_EARLY_EXIT = 1
PDM_USE_GRAPH = 2
PDM_TIME_SOLVES = 4
PDM_NO_AUTOTUNE = 8

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class SceneDescC(C.Structure):
    _fields_ = [
        ("n_vertices", C.c_int), ("x_rest", _dp),
        ("n_tets", C.c_int), ("tets", _ip),
        ("mass_same", C.c_int), ("total_mass", C.c_double),
        ("E", _dp), ("nu_fwd", C.c_double), ("nu_ctor", _dp),
        ("n_attach", C.c_int), ("attach_idx", _ip), ("attach_w", C.c_double),
        ("n_segments", C.c_int), ("seg_start", _dp), ("seg_end", _dp), ("seg_group", _ip),
        ("n_groups", C.c_int), ("k_muscle", C.c_double),
        ("h", C.c_double), ("max_iter", C.c_int), ("damping", C.c_double),
        ("gravity", C.c_double * 3), ("gravity_on", C.c_int),
        ("contact", C.c_int), ("max_iter1", C.c_int), ("alpha", C.c_double), ("mu", C.c_double),
        ("n_planes", C.c_int), ("planes", _dp),
        ("n_spheres", C.c_int), ("spheres", _dp),
        ("nu_fwd_tet", _dp), ("n_bodies", C.c_int), ("body_of_tet", _ip), ("body_mass", _dp),
        ("underwater", C.c_int), ("n_tris", C.c_int), ("tris", _dp),
        ("self_collision", C.c_int), ("self_mu", C.c_double),
    ]


class InfoC(C.Structure):
    _fields_ = [
        ("n_vertices", C.c_int), ("n_tets", C.c_int), ("n_muscles", C.c_int), ("n_groups", C.c_int),
        ("n_supernodes", C.c_int), ("n_levels", C.c_int), ("max_front", C.c_int), ("max_width", C.c_int),
        ("nnz_factor", C.c_longlong), ("nnz_lower", C.c_longlong), ("factor_flops", C.c_double), ("reg_used", C.c_double),
        ("n_contact_vertices", C.c_int), ("batch", C.c_int), ("precision", C.c_int), ("n_drag_faces", C.c_int),
        ("n_obstacle_tris", C.c_int), ("n_self_faces", C.c_int), ("top_rows", C.c_int), ("top_blocks", C.c_int),
        ("leaf_blocks", C.c_int),
    ]


# every symbol declared in include/pdm.h
SYMBOLS = [
    "pdm_last_error", "pdm_device_count", "pdm_scene_create", "pdm_scene_destroy", "pdm_scene_info",
    "pdm_set_stream", "pdm_set_material", "pdm_set_batch", "pdm_forward", "pdm_forward_host", "pdm_backward",
    "pdm_backward_host", "pdm_get_iterations", "pdm_solve_host", "pdm_force_host", "pdm_get_timing",
    "pdm_host_analyse", "pdm_host_solve", "pdm_get_contact_stats", "pdm_time_solve", "pdm_get_solve_timing",
    "pdm_get_solve_shapes", "pdm_set_solve_plan", "pdm_solve_noalloc", "pdm_get_graph_launches",
    "pdm_refactor", "pdm_get_factor_w", "pdm_set_dense_top", "pdm_set_dense_leaves",
]

# PDM_NOALLOC_SCRATCH_BYTES of include/pdm.h
NOALLOC_SCRATCH_BYTES = 64 << 20

_lib = None


def load():
    """Load libpdm.so (raises OSError if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError("libpdm.so not built: run diff_fem_amd.build.build() (hipcc, gfx950)")
    # torch first: its HIP runtime (SONAME libamdhip64.so.7) is then the one
    # libpdm binds to. Loaded the other way round, torch brings a second
    # runtime into the process, which finds no device.
    import torch  # noqa: F401
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.pdm_last_error.restype = C.c_char_p
    L.pdm_device_count.argtypes = [_ip]
    L.pdm_scene_create.argtypes = [C.POINTER(SceneDescC), C.POINTER(vp)]
    L.pdm_scene_destroy.argtypes = [vp]
    L.pdm_scene_destroy.restype = None
    L.pdm_scene_info.argtypes = [vp, C.POINTER(InfoC)]
    L.pdm_set_stream.argtypes = [vp, vp]
    L.pdm_set_material.argtypes = [vp, vp, C.c_double]
    L.pdm_refactor.argtypes = [vp, C.c_int]
    L.pdm_get_factor_w.argtypes = [vp, vp, C.c_longlong]
    L.pdm_set_dense_top.argtypes = [vp, C.c_int]
    L.pdm_set_dense_leaves.argtypes = [vp, C.c_int]
    L.pdm_set_batch.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
    for name in ("pdm_forward", "pdm_forward_host"):
        getattr(L, name).argtypes = [vp, C.c_int, vp, vp, vp, vp, vp, C.c_int]
    for name in ("pdm_backward", "pdm_backward_host"):
        getattr(L, name).argtypes = [vp, vp, vp, vp, vp, vp, vp, vp]
    L.pdm_get_iterations.argtypes = [vp, _ip]
    L.pdm_get_contact_stats.argtypes = [vp, _ip, _ip]
    L.pdm_solve_host.argtypes = [vp, vp, vp]
    L.pdm_solve_noalloc.argtypes = [vp, C.c_int, vp, vp]
    L.pdm_get_graph_launches.argtypes = [vp, C.POINTER(C.c_longlong)]
    L.pdm_force_host.argtypes = [vp, vp, vp, vp]
    L.pdm_get_timing.argtypes = [vp, _dp, _dp, _dp]
    L.pdm_time_solve.argtypes = [vp, C.c_int, _dp]
    L.pdm_host_analyse.argtypes = [C.POINTER(SceneDescC), C.POINTER(InfoC)]
    L.pdm_host_solve.argtypes = [C.POINTER(SceneDescC), C.c_int, vp, vp]
    for s in SYMBOLS:
        if name_missing(L, s):
            raise OSError("libpdm.so lacks symbol " + s)
        if s not in ("pdm_last_error", "pdm_scene_destroy"):
            getattr(L, s).restype = C.c_int
    _lib = L
    return L


def name_missing(L, name):
    try:
        getattr(L, name)
        return False
    except AttributeError:
        return True


class PdmError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        raise PdmError(load().pdm_last_error().decode(errors="replace"))
