"""ctypes binding of libenlightning_hip.so (declared in include/enlightning_hip.h).

The library is the product path: there is no CPU fallback.  Loading fails
with an ImportError-like RuntimeError when the shared object is missing.
"""
import ctypes
import os

import numpy as _np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libenlightning_hip.so")

VARIANTS = [
    "hybrid_high_llf", "hybrid_high", "hybrid_low_llf", "hybrid_low",
    "weno_high_llf", "weno_high", "weno_low_llf", "weno_low", "drp",
]
VARIANT_ID = {n: i for i, n in enumerate(VARIANTS)}
PACK_SINGLE, PACK_STRIDED, PACK_CHUNKED = 0, 1, 2  # ENL_PACK_*
KERNEL_KINDS = ["rhs", "runge_kutta", "update_state", "physical_bc", "copy_interior", "stage_fused"]

# Every exported function and its prototype (argtypes, restype).
_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_RHS_ARGS = [_P] * 27  # all by reference / pointer (enlightningfort.h:23-49)


class EnlGas(ctypes.Structure):
    """Mirror of enl_gas_t."""
    _fields_ = [(n, ctypes.c_double) for n in (
        "weno_alpha", "c", "gamma1", "p0", "T0", "c_lr", "rho0", "R", "c_p", "T_star_n", "T_star_o",
        "mu", "muB", "kappa", "c_v", "R_tilde", "M", "tau_n", "tau_o")] + [
        ("tag_tolerance", ctypes.c_double * 3)]


class EnlSource(ctypes.Structure):
    """Mirror of enl_source_t."""
    _fields_ = [("mode", ctypes.c_int), ("type", ctypes.c_int), ("amp", ctypes.c_double),
                ("spread", ctypes.c_double), ("pos", ctypes.c_double * 2), ("sine_amp", ctypes.c_double),
                ("sine_freq", ctypes.c_double), ("omega", ctypes.c_double), ("nseg", ctypes.c_int),
                ("coords", ctypes.c_void_p)]


class EnlPatchDesc(ctypes.Structure):
    """Mirror of enl_patch_desc_t."""
    _fields_ = [("level", ctypes.c_int), ("box", ctypes.c_int * 4), ("xlo", ctypes.c_double * 2),
                ("dx", ctypes.c_double * 2)]


# enl_patch_desc_t as a numpy record (the ctypes layout above: level and box,
# then the doubles at offset 24)
PATCH_DESC_DTYPE = _np.dtype({"names": ["level", "box", "xlo", "dx"],
                              "formats": [_np.int32, (_np.int32, 4), (_np.float64, 2), (_np.float64, 2)],
                              "offsets": [EnlPatchDesc.level.offset, EnlPatchDesc.box.offset,
                                          EnlPatchDesc.xlo.offset, EnlPatchDesc.dx.offset],
                              "itemsize": ctypes.sizeof(EnlPatchDesc)})

PROTOTYPES = {
    "rhs_": (_RHS_ARGS, None),
    "rhs_hybrid_high_llf_": (_RHS_ARGS, None),
    "rhs_hybrid_high_": (_RHS_ARGS, None),
    "rhs_hybrid_low_llf_": (_RHS_ARGS, None),
    "rhs_hybrid_low_": (_RHS_ARGS, None),
    "rhs_weno_high_llf_": (_RHS_ARGS, None),
    "rhs_weno_high_": (_RHS_ARGS, None),
    "rhs_weno_low_llf_": (_RHS_ARGS, None),
    "rhs_weno_low_": (_RHS_ARGS, None),
    "rhs_drp_": (_RHS_ARGS, None),
    "enl_rhs_variant": ([_I] + _RHS_ARGS + [_P], _I),
    "enl_rhs_variant_noalloc": ([_I] + _RHS_ARGS + [_P], _I),
    "runge_kutta_": ([_P] * 14, None),
    "update_state_": ([_P] * 23, None),
    "tag_cells_": ([_P] * 12, None),
    "enl_set_stream": ([_P], _I),
    "enl_sync": ([], _I),
    "enl_last_error": ([ctypes.c_char_p, ctypes.c_size_t], _I),
    "enl_version": ([], ctypes.c_char_p),
    "enl_fill_physical_bc": ([_I] * 6 + [_P] * 6, _I),
    "enl_profile_start": ([_I], _I),
    "enl_profile_stop": ([_P, _P, _I], _I),
    "enl_fused_scratch_bytes": ([_I] * 5, ctypes.c_size_t),
    "enl_advance_patch_fused": ([_I] * 6 + [_P, _P, ctypes.POINTER(EnlGas), _P, _I, _P, _P, _P, _D, _I, _I]
                                + [_P] * 5 + [ctypes.c_size_t, _D], _I),
    "enl_absorb": ([_I] * 6 + [_P] * 5 + [_D, ctypes.POINTER(EnlGas)] + [_P] * 3, _I),
    "enl_batch_create": ([_I, ctypes.POINTER(EnlPatchDesc), _I, _I, ctypes.POINTER(ctypes.c_void_p)], _I),
    "enl_batch_set_fill": ([_P, _I, _I] + [_P] * 4 + [_I] + [_P] * 6 + [_I] + [_P] * 4 + [_I] + [_P] * 4, _I),
    "enl_batch_set_coarsen": ([_P, _I] + [_P] * 4, _I),
    "enl_batch_set_tiles": ([_P, _I, _I], _I),
    "enl_batch_set_level_tiles": ([_P, _I, _I, _I], _I),
    "enl_batch_set_level_packing": ([_P, _I, _I, _I], _I),
    "enl_switch_create": ([_P], _I),
    "enl_switch_destroy": ([_P], _I),
    "enl_switch_set_params": ([_P, _P], _I),
    "enl_switch_get_params": ([_P, _P], _I),
    "enl_switch_features": ([_I, _I, _I, _I, _I, _P, _P, ctypes.POINTER(EnlGas), _P, _I, _P], _I),
    "enl_switch_forward": ([_P, _P, _I, _P], _I),
    "enl_switch_gradients": ([_P, _P, _P, _P, _I, ctypes.c_float, _P, _P], _I),
    "enl_switch_train_step": ([_P, _P, _P, _P, _I, ctypes.c_float, ctypes.c_float, _P], _I),
    "enl_switch_decide": ([_P, _I, _I, _I, _I, _I, _P, _P, ctypes.POINTER(EnlGas), ctypes.c_float, ctypes.c_float, _P,
                           _P], _I),
    "enl_advance_patch_learned": ([_P, ctypes.c_float, ctypes.c_float, _I, _P, _P, _I, _I, _I, _I, _I, _I, _P, _P,
                                   ctypes.POINTER(EnlGas), _P, _I, _P, _P, _P, ctypes.c_double, _I, _P, _P, _P, _P,
                                   _P, ctypes.c_size_t, _D], _I),
    "enl_batch_set_ambient": ([_P, _I, _I, _I, _P, _P, _P, ctypes.POINTER(EnlGas)], _I),
    "enl_batch_fp32_prepare": ([_P, _I, ctypes.POINTER(EnlGas), _P], _I),
    "enl_batch_fp32_load": ([_P], _I),
    "enl_batch_fp32_set_switch": ([_P, _P, ctypes.c_float, ctypes.c_float, _I, _P], _I),
    "enl_batch_fp32_decisions": ([_P, _I, _P], _I),
    "enl_plan_build": ([_I, _I, _I, _I, _P, _I, _I, _I, _P, _P], _I),
    "enl_plan_from_boxes": ([_I, _I, _I, _I, _P, _P], _I),
    "enl_plan_transfer": ([_P, _P, _I, _P], _I),
    "enl_plan_transfer_rects": ([_P, _P, _I, _P], _I),
    "enl_batch_set_fill_rects": ([_P, _I, _I, _I, _P], _I),
    "enl_batch_set_coarsen_rects": ([_P, _I, _P], _I),
    "enl_batch_transfer_rects": ([_P, _P, _P, _I, _P, _P, _I, _P], _I),
    "enl_batch_tag_map": ([_P, _I, _P, _I, _P, _P, _I], _I),
    "enl_plan_get": ([_P, ctypes.c_char_p, _I, _P, _P], _I),
    "enl_plan_free": ([_P], _I),
    "enl_batch_fp32_store": ([_P], _I),
    "enl_batch_patch_state": ([_P, _I] + [_P] * 4, _I),
    "enl_batch_copy_state": ([_P, _I, _I] + [_P] * 4, _I),
    "enl_batch_prepare": ([_P, _P], _I),
    "enl_batch_zero_source": ([_P], _I),
    "enl_batch_fill": ([_P, _P], _I),
    "enl_batch_update_state": ([_P, ctypes.POINTER(EnlGas)], _I),
    "enl_batch_advance": ([_P, _I, _I, ctypes.POINTER(EnlGas), _P, _I, _P, _P, _P, _D, _I], _I),
    "enl_batch_set_absorb": ([_P, _P, _P, _D], _I),
    "enl_batch_tag_cells": ([_P, _I, _P, _I, _P, _P], _I),
    "enl_batch_transfer": ([_P, _P, _P, _I, _P, _P, _I, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P], _I),
    "enl_add_source": ([_I] * 4 + [_P, _P, ctypes.POINTER(EnlSource), _I, _I, _D, _P, ctypes.c_long], _I),
    "enl_read_lightning": ([ctypes.c_char_p, _P, _I], _I),
    "enl_sample_mics": ([_I] * 4 + [_P, _P, _P, _I, _P, _P, _P, ctypes.c_long, _P, _P], _I),
    "enl_batch_destroy": ([_P], _I),
    "enl_batch_trim": ([], _I),
    "enl_batch_gather": ([_P, _I, _I, _P, ctypes.c_long], _I),
    "enl_batch_scatter": ([_P, _I, _P, ctypes.c_long], _I),
    "enl_advance_patch": ([_I] * 6 + [_P, _P, ctypes.POINTER(EnlGas), _P, _I, _P, _P, _P, _D, _I, _I]
                          + [_P] * 6 + [_D], _I),
}

_lib = None


def load():
    """Load the HIP library (once).  Raises RuntimeError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libenlightning_hip.so is not built (expected at %s); run "
            "`python -c 'import __graft_entry__ as g; g.build()'`" % LIB_PATH)
    # torch first: the library then binds to the HIP runtime torch has
    # loaded, so that device pointers of torch tensors are its own
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    for name, (argtypes, restype) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = restype
    _lib = lib
    return lib


def last_error():
    lib = load()
    buf = ctypes.create_string_buffer(512)
    code = lib.enl_last_error(buf, 512)
    return code, buf.value.decode()


def check(rc=None):
    """Raise if the library recorded an error (rc < 0 or a pending message)."""
    code, msg = last_error()
    if (rc is not None and rc != 0) or code != 0:
        raise RuntimeError("enlightning_hip error %s: %s" % (code if code else rc, msg))
