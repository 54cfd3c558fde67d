"""ctypes binding of librnng_amd.so (the C ABI declared in include/rnng_amd.h).

The library is loaded from this package directory.  There is no fallback: if
the shared object is missing, or no gfx950 device is present when a compute
entry point is called, an exception is raised.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librnng_amd.so")

RNNG_OK = 0
RNNG_ERR_ARG = -1
RNNG_ERR_HIP = -2
RNNG_ERR_CAPACITY = -3
RNNG_ERR_NODEV = -4
RNNG_ERR_FORMAT = -5
RNNG_ERR_UNSUPPORTED = -6

MODE_PLAIN = 0
MODE_MFMA = 1
MODE_GRID = 2


class RnngError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"rnng error {code}: {msg}")
        self.code = code


class Stats(C.Structure):
    """Mirror of rnng_stats (include/rnng_amd.h)."""
    _fields_ = [
        ("pair_evals", C.c_int64),
        ("candidates", C.c_int64),
        ("edges", C.c_int64),
        ("band_pairs", C.c_int64),
        ("band_flips", C.c_int64),
        ("refills", C.c_int64),
        ("prep_ms", C.c_double),
        ("filter_ms", C.c_double),
        ("recheck_ms", C.c_double),
        ("total_ms", C.c_double),
        ("ghost_tests", C.c_int64),
        ("ghost_candidates", C.c_int64),
        ("ghosts", C.c_int64),
        ("ghost_ms", C.c_double),
        ("filter_passes", C.c_int64),
        ("predicted_candidates", C.c_int64),
        ("predicted_edges", C.c_int64),
        ("int8_filter", C.c_int64),
        ("plain_fallbacks", C.c_int64),
        ("call_ms", C.c_double),
        ("host_ms", C.c_double),
        ("kernel_ms", C.c_double),
        ("kernel_units", C.c_int64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


# Every symbol the header declares, with (restype, argtypes).
_vp = C.c_void_p
_i64p = C.POINTER(C.c_int64)
_SIGS = {
    "rnng_version": (C.c_int, []),
    "rnng_last_error": (C.c_char_p, []),
    "rnng_init": (C.c_int, [C.c_int]),
    "rnng_dataset_create_f32": (C.c_int, [_vp, C.c_int64, C.c_int32, C.POINTER(_vp)]),
    "rnng_dataset_wrap_device_f32": (C.c_int, [_vp, C.c_int64, C.c_int32, C.POINTER(_vp)]),
    "rnng_dataset_upload_f32": (C.c_int, [_vp, _vp, C.c_int64]),
    "rnng_dataset_destroy": (C.c_int, [_vp]),
    "rnng_dataset_shape": (C.c_int, [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    "rnng_dataset_copy_f32": (C.c_int, [_vp, _vp]),
    "rnng_dataset_read_file": (C.c_int, [C.c_char_p, C.c_int64, C.c_int64, C.POINTER(_vp)]),
    "rnng_file_info": (C.c_int, [C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "rnng_graph_build": (C.c_int, [_vp, C.c_float, C.c_int32, C.POINTER(_vp), C.POINTER(Stats)]),
    "rnng_graph_build_bipartite": (C.c_int, [_vp, _vp, C.c_float, C.c_int32, C.POINTER(_vp),
                                             C.POINTER(Stats)]),
    "rnng_graph_num_edges": (C.c_int, [_vp, _i64p]),
    "rnng_graph_copy": (C.c_int, [_vp, _vp, _vp, C.c_int64, C.c_int32]),
    "rnng_graph_copy_csr": (C.c_int, [_vp, C.c_int64, C.c_int32, _vp, _vp, _vp, C.c_int64,
                                      C.POINTER(C.c_int64)]),
    "rnng_graph_copy_directed": (C.c_int, [_vp, C.c_int64, _vp, _vp, C.c_int64, _i64p]),
    "rnng_graph_destroy": (C.c_int, [_vp]),
    "rnng_l2_eps_pairs_f32": (C.c_int, [_vp, C.c_int64, C.c_int32, C.c_float, C.c_int32, _vp, _vp,
                                        C.c_int64, _i64p, C.POINTER(Stats)]),
    "rnng_l2_eps_bipartite_f32": (C.c_int, [_vp, C.c_int64, _vp, C.c_int64, C.c_int32, C.c_float,
                                            C.c_int32, _vp, _vp, C.c_int64, _i64p,
                                            C.POINTER(Stats)]),
    "rnng_l2_eps_pairs_f32_noalloc": (C.c_int, [_vp, C.c_int64, C.c_int32, C.c_float, _vp, _vp, C.c_int64,
                                                _i64p, C.POINTER(Stats)]),
    "rnng_set_initial_capacity": (C.c_int, [C.c_int64, C.c_int64]),
    "rnng_rips_build": (C.c_int, [_vp, C.c_float, C.c_int32, C.POINTER(_vp), C.POINTER(Stats)]),
    "rnng_filtration_size": (C.c_int, [_vp, _i64p, _vp]),
    "rnng_filtration_copy": (C.c_int, [_vp, _vp, _vp, C.c_int64]),
    "rnng_filtration_relabel": (C.c_int, [_vp, _vp, C.c_int64, _vp, _vp]),
    "rnng_filtration_destroy": (C.c_int, [_vp]),
    "rnng_strings_create": (C.c_int, [_vp, _vp, C.c_int64, C.POINTER(_vp)]),
    "rnng_strings_read_file": (C.c_int, [C.c_char_p, C.c_int64, C.c_int64, C.POINTER(_vp)]),
    "rnng_strings_size": (C.c_int, [_vp, _i64p, _i64p]),
    "rnng_strings_copy": (C.c_int, [_vp, _vp, _vp]),
    "rnng_strings_destroy": (C.c_int, [_vp]),
    "rnng_edit_graph_build": (C.c_int, [_vp, C.c_float, C.POINTER(_vp), C.POINTER(Stats)]),
    "rnng_edit_graph_build_bipartite": (C.c_int, [_vp, _vp, C.c_float, C.POINTER(_vp), C.POINTER(Stats)]),
    "rnng_edit_eps_pairs": (C.c_int, [_vp, _vp, C.c_int64, C.c_float, _vp, _vp, C.c_int64, _i64p,
                                      C.POINTER(Stats)]),
    "rnng_selection_sample": (C.c_int, [C.c_int64, C.c_int64, _vp, C.c_int32]),
    "rnng_voronoi_assign_f32": (C.c_int, [_vp, C.c_int64, _vp, C.c_int64, C.c_int32, _vp, _vp]),
    "rnng_voronoi_create": (C.c_int, [_vp, _vp, C.c_int64, C.POINTER(_vp)]),
    "rnng_voronoi_create_fps": (C.c_int, [_vp, C.c_int64, C.c_int64, C.POINTER(_vp)]),
    "rnng_voronoi_centers": (C.c_int, [_vp, _vp]),
    "rnng_voronoi_assignment": (C.c_int, [_vp, _vp, _vp]),
    "rnng_voronoi_cells": (C.c_int, [_vp, _vp, _vp]),
    "rnng_voronoi_graph_build": (C.c_int, [_vp, C.c_float, C.c_int32, C.POINTER(_vp),
                                           C.POINTER(Stats)]),
    "rnng_voronoi_set_options": (C.c_int, [_vp, C.c_int32, C.c_int32]),
    "rnng_voronoi_tuning": (C.c_int, [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), _vp]),
    "rnng_voronoi_set_schedule": (C.c_int, [_vp, C.c_int32, C.c_int32]),
    "rnng_voronoi_schedule": (C.c_int, [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), _vp, _vp, C.c_int32,
                                        C.POINTER(C.c_int32)]),
    "rnng_voronoi_num_ghosts": (C.c_int, [_vp, _i64p]),
    "rnng_voronoi_ghosts": (C.c_int, [_vp, _vp, C.c_int64]),
    "rnng_voronoi_destroy": (C.c_int, [_vp]),
    "rnng_mlp_infer": (C.c_int, [_vp, _vp, C.c_int64, _vp]),
    "rnng_mlp_loss_grad": (C.c_int, [_vp, _vp, _vp, C.c_int64, _vp, _vp]),
    "rnng_mlp_train": (C.c_int, [_vp, _vp, _vp, C.c_int64, C.c_int32, C.c_float, _vp]),
    "rnng_set_cost_model": (C.c_int, [_vp]),
    "rnng_mlp_set_under_weight": (C.c_int, [C.c_float]),
    "rnng_mlp_under_weight": (C.c_float, []),
    "rnng_voronoi_cell_stats": (C.c_int, [_vp, _vp, _vp]),
    "rnng_sq_threshold": (C.c_float, [C.c_float]),
    "rnng_debug_mfma_layout": (C.c_int, [_vp, _vp]),
    "rnng_debug_mfma_i8": (C.c_int, [_vp, _vp, _vp]),
    "rnng_set_int8_filter": (C.c_int, [C.c_int32]),
    "rnng_set_filter_kernels": (C.c_int, [C.c_int32, C.c_int32]),
    "rnng_get_filter_kernels": (C.c_int, [C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "rnng_debug_sqrt": (C.c_int, [_vp, _vp, C.c_int64]),
    "rnng_debug_copy": (C.c_int, [C.c_int64, C.POINTER(C.c_double)]),
    "rnng_debug_xcd_permute": (C.c_int, [_vp, C.c_int64, _vp, C.c_int64, _i64p]),
    "rnng_debug_cell_work": (C.c_int, [_vp, C.c_int32, _vp, _vp, C.c_int32, C.c_int32, _vp, _vp, _vp, C.c_int64,
                                       _vp, C.c_int64, _i64p, _i64p, _i64p]),
    "rnng_debug_cell_work_device": (C.c_int, [_vp, C.c_int32, _vp, _vp, C.c_int32, C.c_int32, _vp, _vp, _vp,
                                              C.c_int64, _vp, C.c_int64, _i64p, _i64p, _i64p]),
}

_lib = None


def build(force: bool = False) -> None:
    """Compile the HIP sources for gfx950 with hipcc (make -C csrc)."""
    args = ["make", "-s", "-j8", "-C", os.path.join(_HERE, "csrc")]
    if force:
        subprocess.run(args + ["clean"], check=True)
    subprocess.run(args, check=True)


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def declared_symbols() -> list[str]:
    return sorted(_SIGS)


def check(rc: int) -> None:
    if rc != RNNG_OK:
        msg = lib().rnng_last_error()
        raise RnngError(rc, msg.decode() if msg else "")
