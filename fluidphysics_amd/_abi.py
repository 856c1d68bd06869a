"""ctypes mirror of include/fluidphysics.h and the loader of libfluidphysics.so.

The library is the HIP build under fluidphysics_amd/ (built in-tree by
``__graft_entry__.build()``).  Loading fails loudly when it is missing: there
is no CPU fallback on the product path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfluidphysics.so")

FP_OK = 0
FP_EINVAL = -1
FP_EDEVICE = -2
FP_ENOMEM = -3
FP_ESTATE = -4
FP_EIO = -5


class FpParams(ctypes.Structure):
    """fp_params (include/fluidphysics.h); field order is the ABI."""

    _fields_ = [
        ("gravity", ctypes.c_double),
        ("solid_radius", ctypes.c_double),
        ("fluid_radius", ctypes.c_double),
        ("compliance", ctypes.c_double),
        ("plastic_threshold", ctypes.c_double),
        ("plastic_rate", ctypes.c_double),
        ("break_strain", ctypes.c_double),
        ("fluid_h", ctypes.c_double),
        ("rest_density", ctypes.c_double),
        ("stiffness", ctypes.c_double),
        ("near_stiffness", ctypes.c_double),
        ("viscosity", ctypes.c_double),
        ("solid_damping", ctypes.c_double),
        ("fluid_damping", ctypes.c_double),
        ("floor_friction_solid", ctypes.c_double),
        ("floor_friction_fluid", ctypes.c_double),
        ("wall_x", ctypes.c_double),
        ("wall_z", ctypes.c_double),
        ("wall_height", ctypes.c_double),
        ("wall_half_thick", ctypes.c_double),
        ("cull_radius", ctypes.c_double),
        ("max_speed", ctypes.c_double),
        ("grab_stiffness", ctypes.c_double),
        ("cell_size", ctypes.c_double),
        ("constraint_iters", ctypes.c_int32),
        ("table_size", ctypes.c_int32),
        ("pair_slots", ctypes.c_int32),
        ("fluid_lanes", ctypes.c_int32),
        ("jacobi_omega", ctypes.c_double),
        ("contact_scale", ctypes.c_double),
        ("gravity_x", ctypes.c_double),
        ("gravity_z", ctypes.c_double),
        ("tet_volume", ctypes.c_int32),
        ("solid_contact_cache", ctypes.c_int32),
        ("tear_gap", ctypes.c_double),
    ]

    def to_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}

    def update(self, **kw):
        for k, v in kw.items():
            if k not in dict(self._fields_):
                raise KeyError("unknown solver parameter %r" % k)
            setattr(self, k, v)
        return self


# Reference defaults (src/solver.js:14-37, :50, :220) -- kept here as well so
# the oracle and the scene generators can build params without the library.
REFERENCE_DEFAULTS = dict(
    gravity=-10.0, solid_radius=0.09, fluid_radius=0.075, compliance=2.5e-4,
    plastic_threshold=0.32, plastic_rate=0.04, break_strain=2.6, fluid_h=0.19,
    rest_density=6.0, stiffness=0.002, near_stiffness=0.008, viscosity=0.3,
    solid_damping=1.2, fluid_damping=0.6, floor_friction_solid=0.4,
    floor_friction_fluid=0.04, wall_x=3.5, wall_z=3.5, wall_height=0.92,
    wall_half_thick=0.07, cull_radius=7.0, max_speed=14.0, grab_stiffness=0.3,
    cell_size=0.22, constraint_iters=3, table_size=0,
    pair_slots=0, fluid_lanes=0, jacobi_omega=8.0, contact_scale=1.0, gravity_x=0.0,
    gravity_z=0.0, tet_volume=0, solid_contact_cache=0, tear_gap=1.3,
)


def make_params(**overrides):
    p = FpParams()
    p.update(**REFERENCE_DEFAULTS)
    p.update(**overrides)
    return p


class FpInfo(ctypes.Structure):
    _fields_ = [
        ("capacity", ctypes.c_int32),
        ("solid_count", ctypes.c_int32),
        ("max_fluid", ctypes.c_int32),
        ("table_size", ctypes.c_int32),
        ("num_constraints", ctypes.c_int32),
        ("fluid_next", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("pair_slots", ctypes.c_int32),
        ("substeps_done", ctypes.c_int64),
        ("device_bytes", ctypes.c_int64),
        ("pair_overflow", ctypes.c_int64),
        ("nonfinite", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class FpBatchInfo(ctypes.Structure):
    _fields_ = [
        ("scenes", ctypes.c_int32),
        ("graph", ctypes.c_int32),
        ("capacity", ctypes.c_int64),
        ("device_bytes", ctypes.c_int64),
        ("frames", ctypes.c_int64),
        ("graph_launches", ctypes.c_int64),
        ("contact_lanes", ctypes.c_int32),
        ("graph_captures", ctypes.c_int32),
        ("density_stage", ctypes.c_int32),
        ("xcd_run", ctypes.c_int32),
        ("shape_tuned", ctypes.c_int32),
    ]


class FpPassStat(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 32),
        ("launches", ctypes.c_int64),
        ("total_ms", ctypes.c_double),
    ]


FP_MAX_PASSES = 24

_P = ctypes.c_void_p
_PP = ctypes.POINTER(ctypes.c_void_p)
_D = ctypes.c_double
_I = ctypes.c_int32
_DP = ctypes.POINTER(ctypes.c_double)
_FP = ctypes.POINTER(ctypes.c_float)
_IP = ctypes.POINTER(ctypes.c_int32)
_U8P = ctypes.POINTER(ctypes.c_uint8)

# name -> (restype, argtypes); every symbol include/fluidphysics.h declares.
SIGNATURES = {
    "fp_params_default": (None, [ctypes.POINTER(FpParams)]),
    "fp_create": (_I, [ctypes.POINTER(FpParams), _I, _I, _PP]),
    "fp_destroy": (_I, [_P]),
    "fp_last_error": (ctypes.c_char_p, [_P]),
    "fp_set_params": (_I, [_P, ctypes.POINTER(FpParams)]),
    "fp_get_params": (_I, [_P, ctypes.POINTER(FpParams)]),
    "fp_get_info": (_I, [_P, ctypes.POINTER(FpInfo)]),
    "fp_add_solid_cube": (_I, [_P, _D, _D, _D, _I, _D, _D]),
    "fp_emit": (_I, [_P, _DP, _DP, _DP, _I]),
    "fp_reset_solid": (_I, [_P]),
    "fp_clear_fluid": (_I, [_P]),
    "fp_step": (_I, [_P, _D, _I]),
    "fp_step_async": (_I, [_P, _D, _I]),
    "fp_sync": (_I, [_P]),
    "fp_set_grabs": (_I, [_P, _IP, _DP, _I]),
    "fp_cut": (_I, [_P, _DP, _DP, _DP, _IP]),
    "fp_pick_solid": (_I, [_P, _DP, _DP, _D, _IP, _DP]),
    "fp_select_solids": (_I, [_P, _DP, _D, _I, _IP, _I, _IP]),
    "fp_components": (_I, [_P, _IP, _IP]),
    "fp_fluid_count": (_I, [_P, _IP]),
    "fp_nonfinite": (_I, [_P, _IP]),
    "fp_read_state": (_I, [_P, _FP, _FP, _FP, _FP, _U8P, _IP]),
    "fp_write_state": (_I, [_P, _FP, _FP, _FP, _FP, _U8P]),
    "fp_read_constraints": (_I, [_P, _IP, _IP, _FP, _U8P]),
    "fp_write_constraints": (_I, [_P, _FP, _U8P]),
    "fp_read_tets": (_I, [_P, _IP, _U8P, _IP]),
    "fp_read_cells": (_I, [_P, _I, _DP, _IP, _IP, _FP]),
    "fp_build_grid": (_I, [_P]),
    "fp_read_grid": (_I, [_P, _IP, _IP]),
    "fp_build_grid_noalloc": (_I, [_P, _P, _I, _I, _D, _P, _P, _P]),
    "fp_read_density": (_I, [_P, _FP, _FP]),
    "fp_profile_enable": (_I, [_P, _I]),
    "fp_profile_read": (_I, [_P, ctypes.POINTER(FpPassStat), _I, _IP]),
    "fp_checkpoint_save": (_I, [_P, _I]),
    "fp_checkpoint_restore": (_I, [_P, _I]),
    "fp_state_size": (_I, [_P, ctypes.POINTER(ctypes.c_int64)]),
    "fp_state_export": (_I, [_P, ctypes.c_void_p, ctypes.c_int64]),
    "fp_state_import": (_I, [_P, ctypes.c_void_p, ctypes.c_int64]),
    "fp_batch_create": (_I, [_I, _PP]),
    "fp_batch_destroy": (_I, [_P]),
    "fp_batch_last_error": (ctypes.c_char_p, [_P]),
    "fp_batch_add": (_I, [_P, _P]),
    "fp_batch_remove": (_I, [_P, _P]),
    "fp_batch_step": (_I, [_P, _D, _I]),
    "fp_batch_step_async": (_I, [_P, _D, _I]),
    "fp_batch_sync": (_I, [_P]),
    "fp_batch_set_graph": (_I, [_P, _I]),
    "fp_batch_set_contact_lanes": (_I, [_P, _I]),
    "fp_batch_get_info": (_I, [_P, ctypes.POINTER(FpBatchInfo)]),
    "fp_batch_profile_enable": (_I, [_P, _I]),
    "fp_batch_profile_read": (_I, [_P, ctypes.POINTER(FpPassStat), _I, _IP]),
    "fp_batch_snapshot": (_I, [_P, ctypes.POINTER(_FP), ctypes.POINTER(ctypes.POINTER(ctypes.c_int64)), _IP]),
    "fp_scene_bytes": (_I, [ctypes.POINTER(FpParams), _I, _I, ctypes.POINTER(ctypes.c_int64)]),
    "fp_pool_info": (_I, [_I, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    "fp_pool_trim": (_I, [_I]),
    "fp_tune_cache_save": (_I, [ctypes.c_char_p]),
    "fp_tune_cache_load": (_I, [ctypes.c_char_p]),
    "fp_tune_cache_reset": (_I, []),
}

_lib = None


def load_library(path=None):
    """Load libfluidphysics.so once; raise if it is not built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            "libfluidphysics.so is not built (%s); run __graft_entry__.build()" % p)
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib
