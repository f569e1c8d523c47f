"""ctypes binding of the C ABI (include/openmeeg_amd.h).

The product path has no CPU fallback: importing the binding fails loudly if
libopenmeeg_amd.so is missing, and device entry points need a HIP device.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("OMH_LIBRARY", os.path.join(_HERE, "libopenmeeg_amd.so"))

ORDER_REFERENCE = 0
ORDER_TRIANGLES_FIRST = 1

i32, i64, f64, f32 = C.c_int32, C.c_int64, C.c_double, C.c_float
vp, cp = C.c_void_p, C.c_char_p
p_i64, p_i32, p_f64 = C.POINTER(i64), C.POINTER(i32), C.POINTER(f64)

_SIGS = {
    "omh_last_error": ([], cp),
    "omh_geometry_load": ([cp, cp, C.POINTER(vp)], C.c_int),
    "omh_geometry_nested": ([C.c_int, p_i64, C.POINTER(p_f64), p_i64, C.POINTER(p_i64), p_f64, C.POINTER(vp)], C.c_int),
    "omh_geometry_create": ([C.c_int, C.POINTER(cp), p_i64, C.POINTER(p_f64), p_i64, C.POINTER(p_i64),
                             C.c_int, C.POINTER(cp), p_i64, p_i32, p_i32,
                             C.c_int, C.POINTER(cp), p_f64, p_i64, p_i32, p_i32, C.POINTER(vp)], C.c_int),
    "omh_geometry_free": ([vp], None),
    "omh_geometry_sizes": ([vp, p_i64], C.c_int),
    "omh_geometry_domain_index": ([vp, cp, p_i32], C.c_int),
    "omh_geometry_domain_name": ([vp, i32, C.c_char_p, i64], C.c_int),
    "omh_head_mat": ([vp, C.c_int, vp, i64, C.c_int, vp], C.c_int),
    "omh_head_mat_ex": ([vp, C.c_int, i32, vp, i64, C.c_int, vp], C.c_int),
    "omh_head_mat_integrator": ([vp, C.c_int, C.c_int, f64, vp, i64, C.c_int, vp], C.c_int),
    "omh_locate_domains": ([vp, vp, i64, i64, p_i32, vp], C.c_int),
    "omh_dip_source_mat": ([vp, vp, i64, p_i32, C.c_int, C.c_int, f64, vp, i64, C.c_int, vp], C.c_int),
    "omh_head2eeg_mat": ([vp, vp, i64, p_i64, p_f64, vp], C.c_int),
    "omh_head2meg_mat": ([vp, vp, vp, i64, vp, i64, i32, vp], C.c_int),
    "omh_dip_source2meg_mat": ([vp, i64, vp, vp, i64, vp, i64, vp], C.c_int),
    "omh_mesh_create": ([p_f64, i64, p_i64, i64, C.POINTER(vp)], C.c_int),
    "omh_mesh_load": ([cp, C.POINTER(vp)], C.c_int),
    "omh_surf2vol_mat": ([vp, vp, i64, vp, i64, i32, p_i64, p_i64, vp], C.c_int),
    "omh_eit_source_mat": ([vp, vp, p_f64, i64, C.c_int, vp, i64, i32, vp], C.c_int),
    "omh_dip_source2internal_pot_mat": ([vp, vp, i64, p_i32, vp, i64, vp, i64, p_i64, p_i64, vp], C.c_int),
    "omh_mesh_free": ([vp], None),
    "omh_mesh_sizes": ([vp, p_i64], C.c_int),
    "omh_mesh_intersects": ([vp, vp, p_i32, vp], C.c_int),
    "omh_surf_source_mat": ([vp, vp, C.c_int, C.c_int, vp, i64, i32, vp], C.c_int),
    "omh_surf_source2meg_mat": ([vp, vp, vp, i64, vp, i64, vp], C.c_int),
    "omh_dip_source_adjoint": ([vp, vp, i64, p_i32, C.c_int, vp, i64, C.c_int, vp, vp], C.c_int),
    "omh_geometry_domain_conductivity": ([vp, i32, p_f64], C.c_int),
    "omh_geometry_set_conductivity": ([vp, i32, f64], C.c_int),
    "omh_geometry_pair_count": ([vp, p_i32], C.c_int),
    "omh_geometry_pair_info": ([vp, i32, p_i32, p_i32, p_f64, p_i32], C.c_int),
    "omh_geometry_mesh_info": ([vp, i32, p_i64], C.c_int),
    "omh_batch_create": ([i32, vp, vp, p_i64, vp, p_i64, i32, vp], C.c_int),
    "omh_batch_run": ([vp, i32, p_i32, vp], C.c_int),
    "omh_batch_tune": ([vp, p_f64, i32, p_i32, vp], C.c_int),
    "omh_batch_gain": ([vp, i32, vp, p_i64, p_i64], C.c_int),
    "omh_batch_stats": ([vp, i32, p_i32, p_f64, vp], C.c_int),
    "omh_batch_config": ([vp, p_i32, p_i32, p_i64, p_i64, p_i64, p_i32], C.c_int),
    "omh_batch_free": ([vp], C.c_int),
    "omh_factor_workspace_bytes": ([i64, i32], i64),
    "omh_sym_factor_f64": ([vp, i64, i64, vp, p_i32, vp], C.c_int),
    "omh_sym_factor_f32": ([vp, i64, i64, vp, p_i32, vp], C.c_int),
    "omh_sym_factor_f32_split": ([vp, i64, i64, vp, p_i32, vp], C.c_int),
    "omh_sym_factor_f32_split_blocks": ([vp, i64, i64, vp, i32, vp, vp, p_i32, vp], C.c_int),
    "omh_sym_factor_f32_blocks": ([vp, i64, i64, vp, i32, vp, vp, p_i32, vp], C.c_int),
    "omh_sym_factor_f64_blocks": ([vp, i64, i64, vp, i32, vp, vp, p_i32, vp], C.c_int),
    "omh_sym_solve_f64": ([vp, i64, i64, vp, vp, i64, i64, vp], C.c_int),
    "omh_sym_solve_f32": ([vp, i64, i64, vp, vp, i64, i64, vp], C.c_int),
    "omh_refine_state": ([i32, f64, f64, f64], i32),
    "omh_factor_sign_mismatch": ([vp, i64, i32, i64, p_i64, vp], C.c_int),
    "omh_factor_sign_mismatch_segments": ([vp, i64, i32, i32, vp, vp, p_i64, vp], C.c_int),
    "omh_sym_reorder": ([vp, i64, vp, i64, i32, i64, i32, vp, vp, vp], C.c_int),
    "omh_reorder_rows_f64": ([vp, i64, vp, i64, i64, i64, i32, vp, vp, i32, vp], C.c_int),
    "omh_mixed_solve_reordered": ([vp, i64, i64, vp, i64, vp, vp, i64, vp, i64, i64, i32, f64, i32, vp, vp, vp, p_i32,
                                   p_f64, vp], C.c_int),
    "omh_lu_workspace_bytes": ([i64], i64),
    "omh_lu_factor_f64": ([vp, i64, i64, vp, vp, p_i32, vp], C.c_int),
    "omh_lu_solve_f64": ([vp, i64, i64, vp, vp, vp, i64, i64, vp], C.c_int),
    "omh_mixed_solve": ([vp, i64, i64, vp, i64, vp, vp, i64, vp, i64, i64, i32, f64, p_i32, p_f64, vp], C.c_int),
    "omh_mixed_solve_noalloc": ([vp, i64, i64, vp, i64, vp, vp, i64, vp, i64, i64, i32, f64, p_i32, p_f64, vp],
                                C.c_int),
    "omh_sym_solve_f64_blocks": ([vp, i64, i64, vp, vp, i64, i64, i32, vp, vp, vp], C.c_int),
    "omh_sym_solve_f32_blocks": ([vp, i64, i64, vp, vp, i64, i64, i32, vp, vp, vp], C.c_int),
    "omh_mixed_solve_blocks": ([vp, i64, i64, vp, i64, vp, vp, i64, vp, i64, i64, i32, f64, i32, vp, vp, p_i32, p_f64,
                                vp], C.c_int),
    "omh_gemm_f64": ([i32, i32, i64, i64, i64, f64, vp, i64, vp, i64, f64, vp, i64, vp], C.c_int),
    "omh_gemm_f32": ([i32, i32, i64, i64, i64, f32, vp, i64, vp, i64, f32, vp, i64, vp], C.c_int),
    "omh_convert_f64_f32": ([vp, i64, vp, i64, i64, i64, i32, vp], C.c_int),
    "omh_sparse_rows": ([vp, vp, i64, vp, i64, i64, vp, i64, vp], C.c_int),
    "omh_sparse_transpose_dense": ([vp, vp, i64, i64, vp, i64, vp], C.c_int),
    "omh_sym_unpack": ([vp, i64, i64, vp, i64, C.c_int, vp], C.c_int),
    "omh_sym_pack": ([vp, i64, i64, i64, C.c_int, vp, vp], C.c_int),
    "omh_permute_rows": ([vp, i64, vp, i64, i64, i64, i64, C.c_int, i32, vp], C.c_int),
    "omh_release_scratch": ([], C.c_int),
    "omh_profile_enable": ([i32], C.c_int),
    "omh_profile_read": ([cp, p_i64, p_f64, p_f64], C.c_int),
    "omh_probe_mfma_layout": ([vp, vp], C.c_int),
}

_lib = None


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("openmeeg_amd: %s is missing; build it with `python -m openmeeg_amd.build`" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def symbols():
    return list(_SIGS)


class OpenMEEGError(RuntimeError):
    pass


class OverlappingSourceMesh(OpenMEEGError):
    """OverlappingSourceMesh (OMExceptions.H): a source mesh meets the geometry."""


class NotConverged(OpenMEEGError):
    """Iterative refinement of a mixed-precision solve stopped above its
    tolerance (the results of the last correction are written)."""


STATUS_OVERLAPPING_SOURCE_MESH = 5
STATUS_NOT_CONVERGED = 6


def check(rc):
    if rc == 0:
        return
    msg = lib().omh_last_error().decode()
    if rc == 2:
        raise ValueError(msg)
    if rc == 4:
        raise MemoryError(msg)
    if rc == STATUS_OVERLAPPING_SOURCE_MESH:
        raise OverlappingSourceMesh(msg)
    if rc == STATUS_NOT_CONVERGED:
        raise NotConverged(msg)
    raise OpenMEEGError(msg)
