"""ctypes binding of the C ABI declared in include/patchtable_amd.h.

The shared library patchtable_amd/libpatchtable_hip.so is built in-tree by
patchtable_amd/build.py (hipcc, gfx950). Loading it never falls back to anything else: if the
library is missing, import of the binding raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpatchtable_hip.so")

PT_MAX_DIMS = 64
PT_MFMA_MAX_JUMPS = 16
PT_OK = 0
PT_E_INVALID = -1
PT_BRUTE_BF16 = 0
PT_BRUTE_FP8 = 1


class PtParams(ctypes.Structure):
    """Field-for-field mirror of `pt_params` (include/patchtable_amd.h)."""
    _fields_ = [
        ("randomize_dt", ctypes.c_int),
        ("parallel_dt", ctypes.c_int),
        ("limit", ctypes.c_double),
        ("threads", ctypes.c_int),
        ("dt_threads", ctypes.c_int),
        ("convert_colorspace", ctypes.c_int),
        ("colorspace", ctypes.c_int),
        ("regular_grid", ctypes.c_int),
        ("verbose", ctypes.c_int),
        ("table_knn", ctypes.c_int),
        ("populate_nearest", ctypes.c_int),
        ("populate_random", ctypes.c_int),
        ("partition_step", ctypes.c_int),
        ("check_existing", ctypes.c_int),
        ("do_table_lookup", ctypes.c_int),
        ("init_random", ctypes.c_int),
        ("treecann_eps", ctypes.c_double),
        ("allowed_index", ctypes.c_int),
        ("ndims", ctypes.c_int),
        ("nchroma", ctypes.c_int),
        ("ntables", ctypes.c_int),
        ("nslices_count", ctypes.c_int),
        ("nslices", ctypes.c_int * PT_MAX_DIMS),
        ("populate_exptime", ctypes.c_int),
        ("dt_mode", ctypes.c_int),
        ("dt_algo", ctypes.c_int),
        ("dt_iters", ctypes.c_int),
        ("run_dt", ctypes.c_int),
        ("query_step", ctypes.c_int),
        ("grid_ndims", ctypes.c_int),
        ("prop_dist", ctypes.c_int),
        ("calc_exact_dist", ctypes.c_int),
        ("do_prop", ctypes.c_int),
        ("prop_y_step", ctypes.c_int),
        ("do_rs", ctypes.c_int),
        ("spatial", ctypes.c_int),
        ("prob_rs", ctypes.c_double),
        ("do_calc_dist", ctypes.c_int),
        ("lookup_algo", ctypes.c_int),
        ("is_descriptor", ctypes.c_int),
        ("descriptor_padded", ctypes.c_int),
        ("sanitize_input", ctypes.c_int),
        ("load_filename", ctypes.c_char * 1024),
        ("save_filename", ctypes.c_char * 1024),
        ("patch_w", ctypes.c_int),
        ("coherence_spatial", ctypes.c_double),
        ("coherence_temporal", ctypes.c_double),
        ("recalc_dist_temporal", ctypes.c_int),
        ("table_ratio", ctypes.c_double),
        ("dt_knn", ctypes.c_int),
        ("treecann_agrid", ctypes.c_int),
        ("treecann_bgrid", ctypes.c_int),
        ("flann_build_step", ctypes.c_int),
        ("flann_trees", ctypes.c_int),
        ("flann_checks", ctypes.c_int),
        ("flann_eps", ctypes.c_double),
        ("flann_reset_step", ctypes.c_int),
        ("query_iters", ctypes.c_int),
        ("prop_iters", ctypes.c_int),
        ("enrich_limit", ctypes.c_double),
        ("kcoherence", ctypes.c_int),
        ("kcoherence_min_dist", ctypes.c_int),
        ("kcoherence_algo", ctypes.c_int),
        ("kcoherence_iter", ctypes.c_int),
        ("kcoherence_step", ctypes.c_int),
        ("kcoherence_enrich", ctypes.c_int),
        ("save_kcoherence", ctypes.c_int),
        ("incoherence", ctypes.c_int),
        ("incoherence_min_dist", ctypes.c_int),
        ("triangle_factor", ctypes.c_double),
        ("pm_iters", ctypes.c_int),
        ("pm_knn", ctypes.c_int),
        ("pm_min_dist", ctypes.c_int),
        ("pm_enrich", ctypes.c_int),
        ("pca_samples", ctypes.c_int),
        ("pca_mean", ctypes.c_int),
        ("pca_subtract_mean", ctypes.c_int),
        ("dim_algo", ctypes.c_int),
        ("filter_dims_count", ctypes.c_int),
        ("filter_dims", ctypes.c_int * PT_MAX_DIMS),
        ("auto_filter", ctypes.c_int),
        ("kdtree_add_unique", ctypes.c_int),
        ("ann", ctypes.c_int),
        ("ann_agrid", ctypes.c_int),
        ("ann_bgrid", ctypes.c_int),
        ("ann_eps", ctypes.c_double),
        ("kcoh_window", ctypes.c_int),
        ("jf_rounds", ctypes.c_int),
        ("mfma_rounds", ctypes.c_int),
        ("mfma_anchors", ctypes.c_int),
        ("mfma_jumps", ctypes.c_int * PT_MFMA_MAX_JUMPS),
    ]


_P = ctypes.POINTER
_f32p = _P(ctypes.c_float)
_f64p = _P(ctypes.c_double)
_i32p = _P(ctypes.c_int32)
_i64p = _P(ctypes.c_int64)
_vp = ctypes.c_void_p
_int = ctypes.c_int

# name -> (restype, argtypes)
SIGNATURES = {
    "pt_params_default": (_int, [_P(PtParams)]),
    "pt_params_set_speed": (_int, [_P(PtParams), _int]),
    "pt_params_set_switches": (_int, [_P(PtParams), ctypes.c_char_p]),
    "pt_table_create": (_int, [_P(PtParams), _f32p, _int, _int, _int, _i32p, _P(_vp)]),
    "pt_table_create_dev": (_int, [_P(PtParams), _vp, _int, _int, _int, _i32p, _P(_vp)]),
    "pt_table_lookup": (_int, [_vp, _f32p, _int, _int, _int, _f64p, _f64p, _f32p, _f64p]),
    "pt_table_lookup_dev": (_int, [_vp, _vp, _int, _int, _int, _vp, _vp, _vp, _f64p]),
    "pt_table_lookup_dev_noalloc": (_int, [_vp, _vp, _int, _int, _int, _vp, _vp, _vp, _f64p]),
    "pt_table_save": (_int, [_vp, ctypes.c_char_p]),
    "pt_table_load": (_int, [_P(PtParams), ctypes.c_char_p, _P(_vp)]),
    "pt_table_destroy": (_int, [_vp]),
    "pt_last_error": (ctypes.c_char_p, []),
    "pt_table_info": (_int, [_vp, _P(_int), _P(_int), _i64p, _P(_int), _P(_int)]),
    "pt_table_get_descriptor": (_int, [_vp, _f32p]),
    "pt_table_get_partition": (_int, [_vp, _int, _f32p, _f32p, _f32p, _i32p]),
    "pt_table_get_cells": (_int, [_vp, _i32p]),
    "pt_table_get_populated": (_int, [_vp, _i32p]),
    "pt_table_extra_info": (_int, [_vp, _int, _P(_int), _i64p, _P(_int)]),
    "pt_table_get_extra_cells": (_int, [_vp, _int, _i32p]),
    "pt_table_get_kcoherence": (_int, [_vp, _i32p, _f32p]),
    "pt_table_get_query_descriptor": (_int, [_vp, _f32p]),
    "pt_table_stage_ms": (_int, [_vp, _f64p]),
    "pt_table_set_mfma_stats": (_int, [_vp, _int]),
    "pt_table_mfma_items": (_int, [_vp, ctypes.POINTER(ctypes.c_int64), _int, ctypes.POINTER(_int)]),
    "pt_gck_dev": (_int, [_vp, _int, _int, _int, _int, _int, _vp]),
    "pt_patch_ssd_dev": (_int, [_vp, _int, _vp, _int, _vp, _vp, _vp, _vp, _int, _vp]),
    "pt_desc_dist_dev": (_int, [_vp, _vp, _int, _int, _vp]),
    "pt_synth_pair": (_int, [ctypes.c_uint64, _int, _int, _f32p, _f32p]),
    "pt_table_shape": (_int, [_vp, _P(_int), _P(_int), _P(_int)]),
    "pt_batch_create": (_int, [_P(PtParams), _int, _P(_vp), _P(_int), _P(_int), _int, _P(_vp)]),
    "pt_batch_lookup": (_int, [_vp, _vp, _int, _int, _P(_vp), _f64p, _f64p]),
    "pt_batch_size": (_int, [_vp]),
    "pt_batch_table": (_vp, [_vp, _int]),
    "pt_batch_destroy": (_int, [_vp]),
    "pt_batch_set_tile": (_int, [_vp, _int]),
    "pt_batch_get_tile": (_int, [_vp]),
    "pt_batch_autotune": (_int, [_vp, _vp, _int, _int, _P(_vp), _int, _P(_int), _f64p]),
    "pt_table_set_tile": (_int, [_vp, _int]),
    "pt_table_autotune": (_int, [_vp, _vp, _int, _int, _int, _vp, _int, _P(_int), _f64p]),
    "pt_image_size": (_int, [ctypes.c_char_p, _P(_int), _P(_int), _P(_int)]),
    "pt_image_load": (_int, [ctypes.c_char_p, _f32p, _int, _int, _int]),
    "pt_pfm_save": (_int, [ctypes.c_char_p, _f64p, _int, _int, _int]),
    "pt_key_create": (_int, [ctypes.c_uint64, _P(_vp)]),
    "pt_key_destroy": (_int, [_vp]),
    "pt_key_set_weights": (_int, [_vp, _f32p, _f32p, _f32p, _f32p]),
    "pt_key_get_weights": (_int, [_vp, _f32p, _f32p, _f32p, _f32p, _P(ctypes.c_uint16), _P(ctypes.c_uint16)]),
    "pt_key_forward_dev": (_int, [_vp, _vp, _int, _int, _vp]),
    "pt_key_gradients": (_int, [_vp, _vp, _vp, _int, _int, _vp, _int, ctypes.c_float, _f32p, _f32p, _f32p, _f32p,
                                _f32p]),
    "pt_key_train": (_int, [_vp, _P(PtParams), _int, _P(_vp), _P(_vp), _P(_int), _P(_int), _int, _int, _int,
                            ctypes.c_float, ctypes.c_float, ctypes.c_uint64, _f32p]),
    "pt_key_train_exact": (_int, [_vp, _int, _P(_vp), _P(_vp), _P(_int), _P(_int), _int, _int, _int,
                                  ctypes.c_float, ctypes.c_float, ctypes.c_uint64, _f32p]),
    "pt_table_create_keyed": (_int, [_P(PtParams), _vp, _vp, _int, _int, _P(_vp)]),
    "pt_brute_lookup_dev": (_int, [_vp, _int, _int, _vp, _int, _int, _vp, _int, _int, _vp, _vp, _vp, _vp, _f64p]),
    "pt_brute_exact_f64_dev": (_int, [_vp, _int, _int, _vp, _int, _int, _vp, _int, _vp, _vp]),
}

_lib = None


def load():
    """Loads the in-tree HIP library (raises OSError if it has not been built).

    PyTorch-ROCm ships its own HIP runtime. Two HIP runtimes in one process do not share the
    GPU: the one initialised second reports no device. Importing torch first makes the dynamic
    loader bind the library's libamdhip64.so.7 to torch's already loaded copy, so both use one
    runtime whichever initialises first."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: run patchtable_amd/build.py (or __graft_entry__.build())")
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class PatchTableError(RuntimeError):
    pass


def check(rc):
    if rc != PT_OK:
        msg = load().pt_last_error().decode(errors="replace")
        raise PatchTableError(f"patchtable_amd error {rc}: {msg}")
    return rc
