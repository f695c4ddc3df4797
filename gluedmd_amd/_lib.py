"""ctypes binding of libglued_hip.so (the C-ABI declared in include/glued.h).

The library is the product path: if it has not been built the import of any
compute entry point fails loudly -- there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libglued_hip.so")

c_int_p = ctypes.POINTER(ctypes.c_int)
c_double_p = ctypes.POINTER(ctypes.c_double)
c_float_p = ctypes.POINTER(ctypes.c_float)
c_ll_p = ctypes.POINTER(ctypes.c_longlong)
vp = ctypes.c_void_p
i = ctypes.c_int
d = ctypes.c_double
sz = ctypes.c_size_t

# name -> argtypes (every function returns int status, except glued_last_error)
SIGNATURES = {
    "glued_create": [ctypes.POINTER(vp)],
    "glued_destroy": [vp],
    "glued_set_pbc": [vp, i],
    "glued_set_temperature": [vp, d],
    "glued_add_cv": [vp, i, c_int_p, i, c_double_p, i, c_int_p],
    "glued_add_bias": [vp, i, c_int_p, i, c_double_p, i, c_int_p, i, c_int_p],
    "glued_num_cv_values": [vp, c_int_p],
    "glued_set_test_bias_gradients": [vp, c_double_p, i],
    "glued_share_bias": [vp, i, i],
    "glued_get_multiwalker_ptrs": [vp, i, vp, i, vp],
    "glued_redirect_to_primary_bias": [vp, i, vp, i],
    "glued_initialize": [vp, i, i, c_double_p, i],
    "glued_initialize_batch": [vp, i, c_int_p, c_int_p, c_double_p, i],
    "glued_set_atom_order": [vp, i, c_int_p, i],
    "glued_set_box": [vp, i, c_double_p],
    "glued_set_energy_input": [vp, i, d, vp, vp],
    "glued_execute": [vp, vp, i, vp, vp, i, i, c_double_p],
    "glued_execute_batch": [vp, vp, vp, vp, i, i],
    "glued_execute_plain_all": [vp, vp, vp, vp, i],
    "glued_step_batch": [vp, vp, i, vp, vp, i],
    "glued_update_state": [vp, vp, i],
    "glued_execute_stage": [vp, vp, vp, vp, i],
    "glued_get_cv_values": [vp, i, c_double_p, i],
    "glued_get_bias_energies": [vp, i, c_double_p, i],
    "glued_get_dudcv": [vp, i, c_double_p, i],
    "glued_get_coordination_pairs": [vp, i, i, c_ll_p],
    "glued_get_bias_state": [vp, i, ctypes.c_char_p, sz, ctypes.POINTER(sz)],
    "glued_set_bias_state": [vp, i, ctypes.c_char_p, sz],
    "glued_get_opes_metrics": [vp, i, i, c_double_p],
    "glued_get_kernel_sigmas": [vp, i, i, c_float_p, i, c_int_p],
    "glued_set_hills": [vp, i, i, c_double_p, c_double_p, i],
    "glued_get_hills": [vp, i, i, c_double_p, c_double_p, i, c_int_p],
    "glued_set_grid": [vp, i, i, c_double_p, i],
    "glued_plan_info": [vp, c_int_p, c_int_p],
    "glued_copy_buffer": [vp, i, vp, sz, vp],
    "glued_set_hill_heights": [vp, i, i, vp, i, vp],
    "glued_set_path_lambda": [vp, i, d],
    "glued_set_hill_sigma": [vp, i, c_double_p, i],
    "glued_hill_adjoint": [vp, vp, i, vp, vp, vp],
    "glued_hill_adjoint_noalloc": [vp, vp, i, vp, vp, vp],
    "glued_get_tuning": [vp, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)],
    "glued_opes_adjoint": [vp, vp, i, vp, vp, vp, vp, vp],
    "glued_set_opes_kernels": [vp, i, i, c_double_p, c_double_p, c_double_p, i, ctypes.c_double],
    "glued_set_path_frames": [vp, i, vp, i],
    "glued_path_frames_adjoint": [vp, vp, i, vp, vp],
}

_lib = None


class GluedError(RuntimeError):
    """Raised when a C-ABI call reports failure (the reference raises
    OpenMMException at the same points)."""


def load():
    """Load the built library (raises if it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -m gluedmd_amd.build` (hipcc, gfx950) first"
        )
    lib = ctypes.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    lib.glued_last_error.argtypes = []
    lib.glued_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def call(name, *args):
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        raise GluedError(lib.glued_last_error().decode())
    return rc
