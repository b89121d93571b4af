"""ctypes mirror of include/meshless.h and include/meshless_vis.h.

Structures reproduce the C layouts field by field; every exported function
gets its argument and return types here.  ``lib()`` loads the in-tree
``libmeshless_vis.so`` once and raises ``OSError`` when it is missing: there is
no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_PKG = Path(__file__).resolve().parent


def library_path() -> Path:
    return Path(os.environ.get("MESHLESS_VIS_LIB", _PKG / "libmeshless_vis.so"))


# --------------------------------------------------------------- enums ---
BASIS = {"SPH": 0, "GAUSSIAN": 1, "WENDLAND_D3_C2": 2}
FORM = {"plain": 0, "mfma_f16": 1, "mfma_split_bf16": 2, "mfma_split_f16": 3}
PHI_MODE = {"reference": 0, "analytic": 1}
TILE = {"default": 0, "half_x": 1, "half_y": 2, "quarter": 3}
ERRORS = {
    0: "VIS_SUCCESS",
    1: "VIS_ERROR_INVALID_VALUE",
    2: "VIS_ERROR_INVALID_CONFIG",
    3: "VIS_ERROR_HIP",
    4: "VIS_ERROR_OUT_OF_MEMORY",
    5: "VIS_ERROR_UNSUPPORTED",
    6: "VIS_ERROR_NOT_REGISTERED",
}


# ---------------------------------------------------------- structures ---
class float2(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float)]


class int2(C.Structure):
    _fields_ = [("x", C.c_int), ("y", C.c_int)]


class float3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class Constraint(C.Structure):
    _fields_ = [("position", float3), ("weight", C.c_float)]


class Group(C.Structure):
    _fields_ = [
        ("number_of_terms", C.c_int),
        ("d_number_of_terms", C.c_int),
        ("basis_function_id", C.c_int),
        ("h_constraints", C.POINTER(Constraint)),
        ("d_constraints", C.POINTER(Constraint)),
        ("h_radii", C.POINTER(C.c_float)),
        ("d_radii", C.POINTER(C.c_float)),
    ]


class MeshlessDataset(C.Structure):
    _fields_ = [("groups", C.POINTER(Group)), ("number_of_groups", C.c_int)]


class VisConfig(C.Structure):
    _fields_ = [
        ("step_size", float2),
        ("_cutoff_frequency", int2),
        ("u_axis", float3),
        ("v_axis", float3),
        ("cull_fully_aliased_terms", C.c_bool),
        ("_automatic_d_image", C.c_bool),
        ("_plan", C.c_void_p),
        ("_number_of_samples", int2),
        ("block_length", C.c_int),
        ("_scale", C.c_float),
        ("_d_image", C.c_void_p),
        ("_d_freq_image", C.c_void_p),
        ("_number_of_partial_sums", C.c_int),
        ("_d_freq_image_arranged", C.c_void_p),
        ("form", C.c_int),
        ("phi_mode", C.c_int),
        ("number_of_h_bins", C.c_int),
        ("number_of_interpolation_nodes", C.c_int),
        ("h_bin_ratio", C.c_float),
        ("tile", C.c_int),
    ]


assert C.sizeof(Constraint) == 16
VisConfigP = C.POINTER(VisConfig)
DatasetP = C.POINTER(MeshlessDataset)

# name -> (restype, argtypes); every symbol the headers declare.
PROTOTYPES = {
    # meshless.h
    "get_simple_meshless_dataset": (MeshlessDataset, [C.c_int, C.POINTER(Constraint), C.POINTER(C.c_float), C.c_int]),
    "delete_meshless_dataset": (None, [MeshlessDataset]),
    "delete_meshless_datasets": (None, [DatasetP, C.c_int]),
    "get_number_of_terms": (C.c_int, [DatasetP]),
    "shift_meshless_dataset": (None, [DatasetP, C.c_float, C.c_float, C.c_float]),
    "shift_meshless_datasets": (None, [DatasetP, C.c_int, C.c_float, C.c_float, C.c_float]),
    "load_meshless_datasets_from_file": (None, [C.c_char_p, C.POINTER(DatasetP), C.POINTER(C.c_int)]),
    "save_meshless_datasets_to_file": (None, [C.c_char_p, DatasetP, C.c_int]),
    "meshless_new_constraints": (C.POINTER(Constraint), [C.c_int]),
    "meshless_new_radii": (C.POINTER(C.c_float), [C.c_int]),
    # meshless_vis.h
    "vis_config_create": (VisConfigP, [C.c_bool, float2, int2, float3, float3, int2, C.c_int, C.c_int]),
    "vis_config_get_default": (VisConfigP, []),
    "vis_config_check": (C.c_bool, [VisConfigP]),
    "vis_config_manual_d_image": (None, [VisConfigP, C.c_void_p]),
    "vis_config_change_number_of_samples": (None, [VisConfigP, int2]),
    "vis_config_change_cutoff_frequency": (None, [VisConfigP, int2]),
    "vis_config_change_number_of_partial_sums": (None, [VisConfigP, C.c_int]),
    "vis_config_destroy": (None, [VisConfigP]),
    "vis_config_compute_scale": (None, [VisConfigP]),
    "vis_register_meshless_dataset": (None, [VisConfigP, DatasetP]),
    "vis_unregister_meshless_dataset": (None, [VisConfigP, DatasetP]),
    "vis_opengl_fourier_volume_rendering": (None, [DatasetP, VisConfigP, C.c_uint]),
    "vis_fourier_volume_rendering": (None, [DatasetP, VisConfigP]),
    "vis_copy_to_host": (None, [VisConfigP, C.POINTER(C.c_float)]),
    "fourier_transform": (None, [DatasetP, VisConfigP]),
    "vis_set_stream": (None, [VisConfigP, C.c_void_p]),
    "vis_get_stream": (C.c_void_p, [VisConfigP]),
    "vis_synchronize": (None, [VisConfigP]),
    "vis_copy_spectrum_to_host": (None, [VisConfigP, C.POINTER(float2)]),
    "vis_register_meshless_dataset_async": (None, [VisConfigP, DatasetP]),
    "vis_debug_fail_device_allocations": (None, [C.c_int]),
    "vis_set_partial_budget": (None, [C.c_longlong]),
    "vis_debug_batch_graph_counts": (None, [VisConfigP, C.POINTER(C.c_int), C.POINTER(C.c_longlong)]),
    "vis_fourier_volume_rendering_batched": (None, [C.POINTER(DatasetP), C.POINTER(VisConfigP), C.c_int]),
    "vis_fourier_volume_rendering_noalloc": (None, [DatasetP, VisConfigP]),
    "vis_autotune": (C.c_double, [DatasetP, VisConfigP, C.c_double]),
    "vis_set_kernel_timing": (None, [VisConfigP, C.c_int]),
    "vis_get_kernel_timing": (C.c_int, [VisConfigP, C.POINTER(C.c_double)]),
    "vis_reset_kernel_timing": (None, [VisConfigP]),
    "vis_get_last_error_code": (C.c_int, []),
    "vis_get_last_error": (C.c_char_p, []),
    "vis_clear_last_error": (None, []),
}

_LIB = None


def lib() -> C.CDLL:
    """Load libmeshless_vis.so (once).  Raises OSError when it is not built."""
    global _LIB
    if _LIB is None:
        path = library_path()
        if not path.exists():
            raise OSError(f"{path} not found: build it with `python -m libmeshlessvis_amd.build`")
        L = C.CDLL(str(path))
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


class VisError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code
        self.name = ERRORS.get(code, str(code))


def check(where: str = "") -> None:
    """Raise VisError when the calling thread's last library error is set."""
    L = lib()
    code = L.vis_get_last_error_code()
    if code != 0:
        msg = L.vis_get_last_error().decode(errors="replace")
        L.vis_clear_last_error()
        raise VisError(code, f"{where}: {msg}" if where else msg)
