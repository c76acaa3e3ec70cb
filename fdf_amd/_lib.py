"""ctypes binding of fdf_amd/libpicgpu.so (the C-ABI in include/picgpu.h).

The product path has no CPU fallback: if the gfx950 library is missing this
module raises at import of the binding.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpicgpu.so")

PICGPU_OK = 0
ERRORS = {
    -1: "PICGPU_E_ARG",
    -2: "PICGPU_E_STATE",
    -3: "PICGPU_E_DEGENERATE",
    -4: "PICGPU_E_HIP",
    -5: "PICGPU_E_NOMEM",
    -6: "PICGPU_E_UNSUPPORTED",
    -7: "PICGPU_E_SINGULAR",
    -8: "PICGPU_E_GRID",
}

FP64, FP32 = 0, 1
NEEDS_LOCATION, ACTIVE_AND_LOCATED, MIGRATING_OUT, LOST, UNINITIALIZED = range(5)

FIELD_PARTICLE_COUNT = 0
FIELD_P_AVERAGE = 1
FIELD_CORNER_VELOCITY = 2
FIELD_IFACE, FIELD_JFACE, FIELD_KFACE = 3, 4, 5
FIELD_IFACE_VOLFRAC, FIELD_JFACE_VOLFRAC, FIELD_KFACE_VOLFRAC = 6, 7, 8
FIELD_UCAT, FIELD_UCONT, FIELD_CSI, FIELD_ETA, FIELD_ZET, FIELD_AJ = 9, 10, 11, 12, 13, 14

# boundary faces, types and handlers (common.h:202-222)
FACE_NAMES = ("-Xi", "+Xi", "-Eta", "+Eta", "-Zeta", "+Zeta")
BCTYPE_UNDEFINED, BCTYPE_NOGRAD, BCTYPE_WALL, BCTYPE_SYMMETRY, BCTYPE_INLET, BCTYPE_OUTLET = range(6)
BC_NOGRAD_COPY_GHOST, BC_WALL_NOSLIP, BC_WALL_MOVING, BC_SYMMETRY_PLANE = 1, 2, 3, 4
BC_INLET_PARABOLIC, BC_INLET_CONSTANT_VELOCITY, BC_OUTLET_CONSERVATION = 5, 6, 9

STEP_RK2 = 0x1
STEP_DEPOSIT_CELLS = 0x2
STEP_DEPOSIT_FACES = 0x4
STEP_FIELDS = 0x8
STEP_STAGED = 0x100
STEP_FACES_OF_START = 0x20


class PicgpuOptions(C.Structure):
    _fields_ = [
        ("precision", C.c_int),
        ("use_mlp_guess", C.c_int),
        ("max_traversal", C.c_int),
        ("distance_threshold", C.c_double),
        ("weight_clamp", C.c_double),
        ("bbox_tolerance", C.c_double),
        ("device", C.c_int),
        ("collect_walk_stats", C.c_int),
    ]


class PicgpuStats(C.Structure):
    _fields_ = [
        ("n_particles", C.c_int64),
        ("located", C.c_int64),
        ("lost", C.c_int64),
        ("walk_iterations", C.c_int64),
        ("walk_max", C.c_int64),
        ("fp64_fallbacks", C.c_int64),
        ("removed_lost", C.c_int64),
        ("removed_out_of_bounds", C.c_int64),
        ("particle_steps", C.c_int64),
        ("walk_hist", C.c_int64 * 64),
    ]


class PicgpuBoundaryConfig(C.Structure):
    _fields_ = [
        ("type", C.c_int * 6),
        ("handler", C.c_int * 6),
        ("velocity", (C.c_double * 3) * 6),
        ("vmax", C.c_double * 6),
    ]


class PicgpuError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_P = C.c_void_p
_I64 = C.c_int64
_SIG = {
    "picgpu_default_options": (None, [C.POINTER(PicgpuOptions)]),
    "picgpu_create": (C.c_int, [C.POINTER(_P), C.POINTER(PicgpuOptions)]),
    "picgpu_destroy": (C.c_int, [_P]),
    "picgpu_last_error": (C.c_char_p, []),
    "picgpu_synchronize": (C.c_int, [_P]),
    "picgpu_upload_grid": (C.c_int, [_P, _I64, _I64, _I64, _P]),
    "picgpu_upload_ucat": (C.c_int, [_P, _P]),
    "picgpu_upload_particles": (C.c_int, [_P, _I64, _P, _P, _P, _P, _P, _P, _P]),
    "picgpu_upload_grid_device": (C.c_int, [_P, _I64, _I64, _I64, _P]),
    "picgpu_upload_ucat_device": (C.c_int, [_P, _P]),
    "picgpu_init_particles_device": (C.c_int, [_P, _I64, _P, _I64]),
    "picgpu_num_particles": (C.c_int, [_P, C.POINTER(_I64)]),
    "picgpu_download_particles": (C.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    "picgpu_download_walk_steps": (C.c_int, [_P, _P]),
    "picgpu_reset_statuses": (C.c_int, [_P]),
    "picgpu_advect": (C.c_int, [_P, C.c_double]),
    "picgpu_locate": (C.c_int, [_P]),
    "picgpu_remove_lost": (C.c_int, [_P, C.POINTER(_I64)]),
    "picgpu_remove_out_of_bounds": (C.c_int, [_P, C.POINTER(_I64)]),
    "picgpu_interpolate": (C.c_int, [_P]),
    "picgpu_scatter": (C.c_int, [_P]),
    "picgpu_deposit_faces": (C.c_int, [_P, C.c_double]),
    "picgpu_set_particle_volume": (C.c_int, [_P, C.c_double]),
    "picgpu_step": (C.c_int, [_P, C.c_double, C.c_int, C.c_uint]),
    "picgpu_locate_device_noalloc": (C.c_int, [_P, _I64, _P, _P, _P, _P, C.POINTER(PicgpuStats)]),
    "picgpu_grid_field_size": (C.c_int, [_P, C.c_int, C.POINTER(_I64)]),
    "picgpu_download_grid_field": (C.c_int, [_P, C.c_int, _P]),
    "picgpu_particle_count_summary": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                                C.POINTER(_I64)]),
    "picgpu_interpolation_error": (C.c_int, [_P, C.c_int, C.POINTER(C.c_double)]),
    "picgpu_get_stats": (C.c_int, [_P, C.POINTER(PicgpuStats)]),
    "picgpu_walk_stats_enable": (C.c_int, [_P, C.c_int]),
    "picgpu_timing_enable": (C.c_int, [_P, C.c_int]),
    "picgpu_timing_query": (C.c_int, [_P, C.c_char_p, C.POINTER(C.c_double), C.POINTER(_I64)]),
    "picgpu_timing_reset": (C.c_int, [_P]),
    "picgpu_timing_select": (C.c_int, [_P, C.c_char_p]),
    "picgpu_default_boundary_config": (None, [C.POINTER(PicgpuBoundaryConfig)]),
    "picgpu_parse_boundary_file": (C.c_int, [C.c_char_p, C.POINTER(PicgpuBoundaryConfig)]),
    "picgpu_compute_metrics": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "picgpu_boundary_create": (C.c_int, [_P, C.POINTER(PicgpuBoundaryConfig)]),
    "picgpu_upload_ucont": (C.c_int, [_P, _P]),
    "picgpu_upload_ucont_device": (C.c_int, [_P, _P]),
    "picgpu_set_initial_ucont": (C.c_int, [_P, C.c_int, _P]),
    "picgpu_boundary_execute": (C.c_int, [_P]),
    "picgpu_contra2cart": (C.c_int, [_P]),
    "picgpu_set_eulerian_fields": (C.c_int, [_P, _I64, _I64, C.c_int, _P]),
    "picgpu_read_grid_file": (C.c_int, [C.c_char_p, C.POINTER(_I64), C.POINTER(_I64), C.POINTER(_I64),
                                        C.POINTER(_I64), _P]),
    "picgpu_write_grid_file": (C.c_int, [C.c_char_p, _I64, _I64, _I64, _P]),
    "picgpu_field_file_name": (C.c_int, [C.c_char_p, C.c_size_t, C.c_char_p, C.c_char_p, _I64, C.c_char_p]),
    "picgpu_petsc_vec_length": (C.c_int, [C.c_char_p, C.POINTER(_I64)]),
    "picgpu_read_petsc_vec": (C.c_int, [C.c_char_p, _I64, _P]),
    "picgpu_write_petsc_vec": (C.c_int, [C.c_char_p, _I64, C.c_int, _P]),
    "picgpu_write_output": (C.c_int, [_P, C.c_char_p, _I64]),
    "picgpu_read_restart": (C.c_int, [_P, C.c_char_p, _I64]),
    "picgpu_write_vts": (C.c_int, [C.c_char_p, _I64, _I64, _I64, _P, C.c_char_p, C.c_int, _P]),
    "picgpu_write_vtp": (C.c_int, [C.c_char_p, _I64, _P, C.c_char_p, C.c_int, _P]),
    "picgpu_write_vtk": (C.c_int, [_P, C.c_char_p, _I64]),
    "picgpu_seed_particles": (C.c_int, [_P, _I64, C.c_int, C.c_int, _I64, _I64]),
    "picgpu_reinit_on_inlet": (C.c_int, [_P, C.c_int, _I64]),
    "picgpu_mlp_param_count": (_I64, []),
    "picgpu_mlp_set_params": (C.c_int, [_P, _P]),
    "picgpu_mlp_get_params": (C.c_int, [_P, _P]),
    "picgpu_mlp_train": (C.c_int, [_P, _I64, C.c_int, C.c_int, C.c_float, C.c_uint64, C.POINTER(C.c_float)]),
    "picgpu_mlp_forward": (C.c_int, [_P, _I64, _P, _P]),
    "picgpu_mlp_loss_grad": (C.c_int, [_P, _I64, _P, _P, C.POINTER(C.c_float), _P]),
    "picgpu_mlp_guess": (C.c_int, [_P, _I64, _P, _P]),
    "picgpu_mlp_enable": (C.c_int, [_P, C.c_int]),
}

_lib = None


def exported_symbols() -> list[str]:
    return sorted(_SIG)


def load() -> C.CDLL:
    """Load libpicgpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
        )
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIG.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code: int) -> None:
    if code != PICGPU_OK:
        msg = load().picgpu_last_error()
        raise PicgpuError(code, msg.decode() if msg else "")
