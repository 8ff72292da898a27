"""ctypes binding of the C ABI in include/rh264_me.h (librh264_me.so).

The library is the product: every function here calls a HIP kernel on the
GPU.  There is no CPU fallback -- if the shared library is missing or no GPU is
present, calls raise.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# RH264_LIB selects another build of the same library (e.g. a tuning variant)
LIB_PATH = os.environ.get("RH264_LIB") or os.path.join(_HERE, "librh264_me.so")
LR_N = 128

c_u8p = C.POINTER(C.c_uint8)
c_i32p = C.POINTER(C.c_int32)
c_i64p = C.POINTER(C.c_int64)
c_vp = C.c_void_p


class SearchParams(C.Structure):
    """rh264_search_params (include/rh264_me.h)."""
    _fields_ = [
        ("cw", C.c_int32), ("ch", C.c_int32), ("pad", C.c_int32),
        ("bw", C.c_int32), ("bh", C.c_int32), ("range", C.c_int32),
        ("subpel_pat", C.c_int32),
        ("lr_fp", C.c_int32 * LR_N), ("lr", C.c_int32 * LR_N), ("hadamard", C.c_int32),
    ]


class Diamond(C.Structure):
    """rh264_diamond (include/rh264_me.h)."""
    _fields_ = [("steps", C.c_int32 * 8), ("nsteps", C.c_int32), ("sadfp", C.c_int32), ("snap", C.c_int32),
                ("fc", C.c_int32), ("sad_only", C.c_int32), ("mv_cost", C.c_int32), ("rescue", C.c_int32),
                ("rescue_range", C.c_int32), ("rescue_thresh", C.c_int32), ("rescue_fast", C.c_int32),
                ("rescue_var", C.c_int32), ("lam", C.c_double), ("lam_fp", C.c_double), ("mv_tab", C.c_void_p),
                ("rescue_learn", C.c_int32), ("rescue_payoff", C.c_int32)]


class Stream(C.Structure):
    """rh264_stream (include/rh264_me.h)."""
    _fields_ = [("cur", C.c_void_p), ("planes", C.c_void_p), ("mvps", C.c_void_p), ("cw", C.c_int32),
                ("ch", C.c_int32), ("nrefs", C.c_int32), ("out", C.c_void_p)]


class Seq(C.Structure):
    """rh264_seq (include/rh264_me.h)."""
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("qp", C.c_int32), ("num_ref_frames", C.c_int32),
                ("level_idc", C.c_int32)]


# name -> (restype, argtypes); mirrors include/rh264_me.h one to one
PROTOTYPES = {
    "rh264_write_headers": (C.c_int, [C.POINTER(Seq), c_vp, C.c_size_t, C.POINTER(C.c_size_t)]),
    "rh264_write_idr_pcm": (C.c_int, [C.POINTER(Seq), c_vp, c_vp, c_vp, c_vp, C.c_size_t, C.POINTER(C.c_size_t)]),
    "rh264_write_p_slice": (C.c_int, [C.POINTER(Seq), C.c_int, C.c_int, C.c_int, c_vp, c_vp, C.c_size_t,
                                      C.POINTER(C.c_size_t)]),
    "rh264_write_p_slice_residual": (C.c_int, [C.POINTER(Seq), C.c_int, C.c_int, C.c_int, c_vp, c_vp, c_vp, c_vp,
                                               c_vp, C.c_size_t, C.POINTER(C.c_size_t)]),
    "rh264_init": (C.c_int, [C.c_int, C.POINTER(c_vp)]),
    "rh264_destroy": (C.c_int, [c_vp]),
    "rh264_last_error": (C.c_char_p, []),
    "rh264_device_name": (C.c_int, [c_vp, C.c_char_p, C.c_int]),
    "rh264_alloc": (C.c_int, [c_vp, C.c_size_t, C.POINTER(c_vp)]),
    "rh264_free": (C.c_int, [c_vp, c_vp]),
    "rh264_h2d": (C.c_int, [c_vp, c_vp, c_vp, C.c_size_t]),
    "rh264_d2h": (C.c_int, [c_vp, c_vp, c_vp, C.c_size_t]),
    "rh264_memset": (C.c_int, [c_vp, c_vp, C.c_int, C.c_size_t]),
    "rh264_sync": (C.c_int, [c_vp]),
    "rh264_event_record": (C.c_int, [c_vp, C.c_int]),
    "rh264_event_elapsed_ms": (C.c_int, [c_vp, C.c_int, C.c_int, C.POINTER(C.c_float)]),
    "rh264_block_metric": (C.c_int, [c_vp, C.c_int, C.c_int, C.c_int, C.c_int, c_vp, C.c_int, c_vp,
                                     c_vp, C.c_int, c_vp, c_vp, c_vp]),
    **{f"rh264_{k}_{s}": (C.c_int, [c_vp, C.c_size_t, c_vp, C.c_size_t, c_vp])
       for k, shapes in (("satd", ("16x16", "16x8", "8x16", "8x8", "4x4")), ("sad", ("16x16", "16x8", "8x16")))
       for s in shapes},
    "rh264_sad_x4": (C.c_int, [c_vp, C.c_size_t, c_vp, c_vp, C.c_size_t, C.c_size_t, C.c_size_t, c_vp]),
    "rh264_satd_x4": (C.c_int, [c_vp, C.c_size_t, c_vp, c_vp, C.c_size_t, C.c_size_t, C.c_size_t, c_vp]),
    "rh264_satd_x4p": (C.c_int, [c_vp, C.c_size_t, c_vp, c_vp, C.c_size_t, C.c_size_t, C.c_size_t, c_vp]),
    "rh264_satd_avg_x4": (C.c_int, [c_vp, C.c_size_t, c_vp, c_vp, c_vp, c_vp, C.c_size_t, C.c_size_t, C.c_size_t,
                                    c_vp]),
    "rh264_satd_avg": (C.c_int, [c_vp, C.c_size_t, c_vp, c_vp, C.c_size_t, C.c_size_t, C.c_size_t, c_vp]),
    "rh264_hpel_fused": (C.c_int, [c_vp, C.c_size_t, C.c_size_t, c_vp, c_vp, c_vp]),
    "rh264_mectx_new": (C.c_int, [c_vp, C.c_size_t, C.c_size_t, c_vp, c_vp, c_vp, c_vp, C.c_size_t, C.c_size_t,
                                  C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t,
                                  C.POINTER(c_vp)]),
    "rh264_mectx_eval": (C.c_int, [c_vp, C.c_int32, C.c_int32, C.POINTER(C.c_uint32)]),
    "rh264_mectx_eval_n": (C.c_int, [c_vp, C.c_int, c_vp, c_vp, c_vp]),
    "rh264_mectx_free": (C.c_int, [c_vp]),
    "rh264_accel_set_device": (C.c_int, [C.c_int]),
    "rh264_accel_release": (C.c_int, []),
    "rh264_planes_bytes": (C.c_size_t, [C.c_int, C.c_int, C.c_int]),
    "rh264_build_planes": (C.c_int, [c_vp, c_vp, C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_int, c_vp]),
    "rh264_lambda_table": (None, [C.c_double, c_i32p]),
    "rh264_full_search": (C.c_int, [c_vp, C.POINTER(SearchParams), C.c_int, c_vp, C.c_size_t, c_vp,
                                    C.c_size_t, c_vp, c_vp, c_vp, c_vp]),
    "rh264_subpel_refine": (C.c_int, [c_vp, C.POINTER(SearchParams), C.c_int, c_vp, C.c_size_t, c_vp,
                                      C.c_size_t, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "rh264_plain_search": (C.c_int, [c_vp, C.POINTER(SearchParams), C.c_int, c_vp, C.c_size_t, c_vp,
                                     C.c_size_t, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "rh264_set_profiling": (C.c_int, [c_vp, C.c_int]),
    "rh264_profile_count": (C.c_int, [c_vp]),
    "rh264_profile_get": (C.c_int, [c_vp, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "rh264_profile_phases": (C.c_int, [c_vp, C.c_int, C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_int)]),
    "rh264_profile_times": (C.c_int, [c_vp, C.c_int, C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_int)]),
    "rh264_tune": (C.c_int, [c_vp, C.c_char_p, C.c_int]),
    "rh264_mbtree_downsample": (C.c_int, [c_vp, c_vp, C.c_size_t, C.c_int, C.c_int, C.c_int, c_vp]),
    "rh264_mbtree_costs": (C.c_int, [c_vp, c_vp, C.c_int, C.c_int, C.c_int, c_vp, c_vp, c_vp]),
    "rh264_mbtree_propagate_i32": (C.c_int, [c_vp, C.c_int, C.c_int, C.c_int, c_vp, c_vp, c_vp, c_vp]),
    "rh264_mbtree_costs_multi": (C.c_int, [c_vp, c_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, c_vp, c_vp, c_vp,
                                           c_vp]),
    "rh264_mbtree_costs_multi_learned": (C.c_int, [c_vp, c_vp, c_vp, C.c_size_t, c_vp, C.c_int, C.c_int, C.c_int,
                                                   C.c_int, C.c_int, C.c_int, C.c_int32, c_vp, c_vp, c_vp, c_vp,
                                                   c_vp]),
    "rh264_mbtree_propagate_multi_i32": (C.c_int, [c_vp, C.c_int, C.c_int, C.c_int, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "rh264_plain_search_noalloc": (C.c_int, [c_vp, C.POINTER(SearchParams), C.c_int, c_vp, C.c_size_t, c_vp,
                                             C.c_size_t, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "rh264_mv_cost_smooth": (None, [C.POINTER(C.c_uint16)]),
    "rh264_motion_search": (C.c_int, [c_vp, C.POINTER(SearchParams), C.POINTER(Diamond), C.c_int, c_vp, C.c_size_t,
                                      c_vp, C.c_size_t, c_vp, C.c_int, c_vp, c_vp, c_vp]),
    "rh264_motion_search_wavefront": (C.c_int, [c_vp, C.POINTER(SearchParams), C.POINTER(Diamond), C.c_int,
                                                C.c_int, C.c_double, c_vp, C.c_size_t, c_vp, C.c_size_t, c_vp, c_vp,
                                                c_vp]),
    "rh264_partition_decision": (C.c_int, [c_vp, C.POINTER(SearchParams), C.POINTER(Diamond), C.c_int, C.c_int,
                                           C.c_double, C.c_int32, C.c_int, c_vp, C.c_size_t, c_vp, C.c_size_t,
                                           c_vp]),
    "rh264_split_gate": (C.c_int32, [C.c_int]),
    "rh264_decision_predict": (C.c_int, [c_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, c_vp, C.c_size_t, c_vp,
                                         c_vp, c_vp, c_vp, c_vp]),
    "rh264_mbtree_qp_offsets": (C.c_int, [C.c_int, C.c_int, C.c_int, c_vp, c_vp, c_vp, C.c_double, C.c_double,
                                         C.c_double, C.c_int, c_vp, c_vp]),
    "rh264_grain_signature": (C.c_int, [c_vp, c_vp, C.c_int, C.c_int, C.POINTER(C.c_int32), c_vp]),
    "rh264_rc_complexity": (C.c_int, [c_vp, c_vp, c_vp, C.c_size_t, C.c_int, C.c_int, C.c_int, c_vp]),
    "rh264_batch_out_len": (C.c_size_t, [C.c_int, C.c_int, c_i32p, C.c_int]),
    "rh264_batch_mvp_len": (C.c_size_t, [C.c_int, C.c_int, c_i32p, C.c_int]),
    "rh264_batch_create": (C.c_int, [c_vp, C.POINTER(Stream), C.c_int, c_i32p, C.c_int, C.c_int, C.c_int, C.c_int,
                                     c_i32p, c_i32p, C.c_double, C.POINTER(c_vp)]),
    "rh264_batch_run": (C.c_int, [c_vp, c_vp]),
    "rh264_batch_search": (C.c_int, [c_vp, c_vp]),
    "rh264_batch_autotune": (C.c_int, [c_vp, c_vp, C.c_int]),
    "rh264_batch_choice": (C.c_int, [c_vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "rh264_batch_capture": (C.c_int, [c_vp, c_vp]),
    "rh264_batch_replay": (C.c_int, [c_vp, c_vp]),
    "rh264_batch_destroy": (C.c_int, [c_vp]),
    "rh264_mvnet_param_count": (C.c_size_t, []),
    "rh264_mvnet_create": (C.c_int, [c_vp, C.c_int, C.c_int, C.c_int, C.POINTER(c_vp)]),
    "rh264_mvnet_destroy": (C.c_int, [c_vp]),
    "rh264_mvnet_set_params": (C.c_int, [c_vp, c_vp, C.POINTER(C.c_float), C.c_size_t]),
    "rh264_mvnet_get_params": (C.c_int, [c_vp, c_vp, C.POINTER(C.c_float), C.c_size_t]),
    "rh264_mvnet_get_grads": (C.c_int, [c_vp, c_vp, C.POINTER(C.c_float), C.c_size_t]),
    "rh264_mvnet_forward": (C.c_int, [c_vp, c_vp, c_vp, c_vp, C.c_size_t, C.c_int, C.c_int, C.c_int, c_vp, c_vp]),
    "rh264_mvnet_backward": (C.c_int, [c_vp, c_vp, c_vp, c_vp, C.c_size_t, C.c_int, C.c_int, C.c_int, c_vp,
                                       C.POINTER(C.c_float)]),
    "rh264_mvnet_adam": (C.c_int, [c_vp, c_vp, C.c_float]),
    "rh264_learned_search": (C.c_int, [c_vp, C.POINTER(SearchParams), c_vp, C.c_int, c_vp, c_vp, C.c_size_t, c_vp,
                                       C.c_size_t, C.c_int, C.c_int32, c_vp, c_vp, c_vp]),
    "rh264_predict_blocks": (C.c_int, [c_vp, c_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                       c_vp, c_vp, c_vp]),
}

_lib = None


def load():
    """Load librh264_me.so (built by __graft_entry__.build()); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                           "(the HIP provider has no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
