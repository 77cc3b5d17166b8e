"""ctypes mirror of include/hcx.h (the libhcx C-ABI)."""
import ctypes as C
import os

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "libhcx.so")


class HcxComplex(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class TrackOpts(C.Structure):
    _fields_ = [
        ("max_steps", C.c_int), ("max_corr", C.c_int), ("incr_steps", C.c_int),
        ("dt0", C.c_double), ("end_zone", C.c_double), ("t_tol", C.c_double), ("dt_min", C.c_double),
        ("corr_tol", C.c_double), ("inf_tol", C.c_double),
        ("refine_iters", C.c_int), ("refine_tol", C.c_double), ("real_tol", C.c_double),
        ("unique_tol", C.c_double), ("precision", C.c_int), ("lanes_per_path", C.c_int),
        ("waves_per_cu", C.c_int), ("specialize", C.c_int), ("retrack_duplicates", C.c_int), ("fp32_corrector", C.c_int),
        ("real_ref_tol", C.c_double), ("replicas", C.c_int),
    ]


class ProblemDesc(C.Structure):
    _fields_ = [
        ("num_vars", C.c_int), ("num_coeffs", C.c_int), ("max_order_t", C.c_int),
        ("hx_terms", C.c_int), ("hx_parts", C.c_int), ("ht_terms", C.c_int), ("ht_parts", C.c_int),
        ("h_hx_idx", C.POINTER(C.c_int)), ("h_ht_idx", C.POINTER(C.c_int)),
        ("num_tracks", C.c_int), ("h_start_sols", C.c_void_p),
        ("num_params", C.c_int), ("h_coeff_affine", C.c_void_p), ("h_start_params", C.c_void_p),
        ("cmap_terms", C.c_int), ("cmap_deg", C.c_int), ("h_cmap_scalar", C.c_void_p),
        ("h_cmap_params", C.POINTER(C.c_int)),
    ]


class ProblemInfo(C.Structure):
    _fields_ = [
        ("num_vars", C.c_int), ("num_coeffs", C.c_int), ("max_order_t", C.c_int), ("num_tracks", C.c_int),
        ("num_params", C.c_int), ("hx_terms", C.c_int), ("ht_terms", C.c_int), ("hx_vars", C.c_int),
        ("ht_vars", C.c_int),
        ("flops_coef", C.c_double), ("flops_hx", C.c_double), ("flops_ht", C.c_double), ("flops_h", C.c_double),
        ("flops_lu", C.c_double), ("flops_stage", C.c_double), ("flops_newton", C.c_double),
    ]


class BatchDesc(C.Structure):
    _fields_ = [
        ("num_instances", C.c_int),
        ("d_coeffs_hx", C.c_void_p), ("d_coeffs_ht", C.c_void_p),
        ("d_target_params", C.c_void_p), ("d_start_params", C.c_void_p), ("d_coeff_work", C.c_void_p),
    ]


class ResultDesc(C.Structure):
    _fields_ = [
        ("d_roots", C.c_void_p), ("d_flags", C.c_void_p), ("d_steps", C.c_void_p), ("d_newton", C.c_void_p),
        ("d_refine", C.c_void_p), ("d_t_final", C.c_void_p), ("d_dt_final", C.c_void_p),
        ("d_residual", C.c_void_p), ("d_inst_done", C.c_void_p), ("d_batch_start", C.c_void_p),
    ]


class Counts(C.Structure):
    _fields_ = [("tracked", C.c_int64), ("infinity", C.c_int64), ("refined", C.c_int64), ("real", C.c_int64),
                ("unique", C.c_int64), ("unique_ref", C.c_int64), ("real_ref", C.c_int64)]


class SynthSpec(C.Structure):
    _fields_ = [("num_vars", C.c_int), ("num_params", C.c_int), ("num_quadratic", C.c_int),
                ("monomials", C.c_int), ("seed", C.c_uint64)]


class SystemInfo(C.Structure):
    _fields_ = [("num_vars", C.c_int), ("num_params", C.c_int), ("num_coeffs", C.c_int),
                ("max_degree", C.c_int), ("max_terms", C.c_int),
                ("hx_terms", C.c_int), ("hx_parts", C.c_int), ("ht_terms", C.c_int), ("ht_parts", C.c_int)]


class RefSettings(C.Structure):
    _fields_ = [("problem_name", C.c_char * 128), ("gpuhc_type", C.c_char * 16),
                ("max_steps", C.c_int), ("max_corr", C.c_int), ("incr_steps", C.c_int),
                ("num_vars", C.c_int), ("num_params", C.c_int), ("num_tracks", C.c_int),
                ("hx_terms", C.c_int), ("hx_parts", C.c_int), ("ht_terms", C.c_int), ("ht_parts", C.c_int),
                ("max_order_t", C.c_int), ("num_coeffs", C.c_int)]


class TuneResult(C.Structure):
    _fields_ = [("lanes_per_path", C.c_int), ("specialize", C.c_int), ("waves_per_cu", C.c_int),
                ("best_ms", C.c_double), ("candidates", C.c_int)]


class ServerOpts(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_int64), ("max_batch_requests", C.c_int), ("streams", C.c_int),
                ("max_queue", C.c_int)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("rejected", C.c_int64),
                ("inflight_requests", C.c_int64), ("inflight_paths", C.c_int64), ("launches", C.c_int64),
                ("paths_done", C.c_int64), ("converged_paths", C.c_int64), ("hbm_in_use", C.c_int64),
                ("peak_inflight_paths", C.c_int64), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
                ("mean_ms", C.c_double), ("max_ms", C.c_double)]


class RequestResult(C.Structure):
    _fields_ = [("type_id", C.c_int), ("num_paths", C.c_int), ("refined", C.c_int), ("real", C.c_int),
                ("latency_ms", C.c_double), ("queue_ms", C.c_double)]


# name -> (restype, argtypes).  Every symbol here is declared in include/hcx.h.
SIGNATURES = {
    "hcx_last_error": (C.c_char_p, []),
    "hcx_abi_version": (C.c_int, []),
    "hcx_device_count": (C.c_int, []),
    "hcx_track_opts_default": (None, [C.POINTER(TrackOpts)]),
    "hcx_problem_create": (C.c_int, [C.POINTER(ProblemDesc), C.POINTER(C.c_void_p)]),
    "hcx_problem_destroy": (C.c_int, [C.c_void_p]),
    "hcx_problem_get_info": (C.c_int, [C.c_void_p, C.POINTER(ProblemInfo)]),
    "hcx_track": (C.c_int, [C.c_void_p, C.POINTER(BatchDesc), C.POINTER(TrackOpts), C.POINTER(ResultDesc),
                            C.c_void_p]),
    "hcx_track_static": (C.c_int, [C.c_void_p, C.POINTER(BatchDesc), C.POINTER(TrackOpts), C.POINTER(ResultDesc),
                                   C.c_void_p]),
    "hcx_track_noalloc": (C.c_int, [C.c_void_p, C.POINTER(BatchDesc), C.POINTER(TrackOpts), C.POINTER(ResultDesc),
                                    C.c_void_p]),
    "hcx_track_multi": (C.c_int, [C.c_int, C.POINTER(C.c_void_p), C.POINTER(BatchDesc), C.POINTER(ResultDesc),
                                  C.POINTER(TrackOpts), C.c_void_p]),
    "hcx_kernel_HC_Solver": (C.c_double, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p]),
    "hcx_sync": (C.c_int, [C.c_void_p]),
    "hcx_stream_release": (C.c_int, [C.c_void_p]),
    "hcx_autotune": (C.c_int, [C.c_void_p, C.POINTER(BatchDesc), C.POINTER(TrackOpts), C.POINTER(TuneResult),
                               C.c_void_p]),
    "hcx_problem_set_launch_shape": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "hcx_graph_capture_multi": (C.c_int, [C.c_int, C.POINTER(C.c_void_p), C.POINTER(BatchDesc),
                                          C.POINTER(ResultDesc), C.POINTER(TrackOpts), C.POINTER(C.c_void_p)]),
    "hcx_graph_launch": (C.c_int, [C.c_void_p, C.c_void_p]),
    "hcx_graph_destroy": (C.c_int, [C.c_void_p]),
    "hcx_server_opts_default": (None, [C.POINTER(ServerOpts)]),
    "hcx_server_create": (C.c_int, [C.POINTER(ServerOpts), C.POINTER(TrackOpts), C.POINTER(C.c_void_p)]),
    "hcx_server_add_type": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]),
    "hcx_server_submit": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.POINTER(C.c_uint64)]),
    "hcx_server_poll": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_int), C.POINTER(RequestResult), C.c_void_p,
                                  C.c_void_p]),
    "hcx_server_drain": (C.c_int, [C.c_void_p, C.c_double]),
    "hcx_server_get_stats": (C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
    "hcx_server_reset_stats": (C.c_int, [C.c_void_p]),
    "hcx_server_destroy": (C.c_int, [C.c_void_p]),
    "hcx_collect": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_double, C.POINTER(Counts),
                              C.c_void_p]),
    "hcx_collect_host": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_double,
                                   C.POINTER(Counts)]),
    "hcx_system_create_synthetic": (C.c_int, [C.POINTER(SynthSpec), C.POINTER(C.c_void_p)]),
    "hcx_system_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.POINTER(C.c_void_p)]),
    "hcx_system_destroy": (C.c_int, [C.c_void_p]),
    "hcx_system_get_info": (C.c_int, [C.c_void_p, C.POINTER(SystemInfo)]),
    "hcx_system_get_monomials": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "hcx_system_get_affine": (C.c_int, [C.c_void_p, C.c_void_p]),
    "hcx_system_compile_tables": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "hcx_system_eval_coeffs": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "hcx_system_param_homotopy_coeffs": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "hcx_system_problem_desc": (C.c_int, [C.c_void_p, C.POINTER(ProblemDesc)]),
    "hcx_total_degree_homotopy": (C.c_int, [C.c_void_p, C.c_void_p, HcxComplex, C.POINTER(C.c_void_p),
                                            C.POINTER(C.c_int), C.c_void_p, C.c_void_p]),
    "hcx_total_degree_start_roots": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "hcx_total_degree_subset": (C.c_int, [C.c_void_p, C.c_int, C.c_uint64, C.c_void_p]),
    "hcx_random_cn": (C.c_int, [C.c_uint64, C.c_int64, C.c_void_p]),
    "hcx_random_gamma": (HcxComplex, [C.c_uint64]),
    "hcx_ref_read_settings": (C.c_int, [C.c_char_p, C.POINTER(RefSettings)]),
    "hcx_ref_read_params": (C.c_int, [C.c_char_p, C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    "hcx_ref_read_start_sols": (C.c_int, [C.c_char_p, C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    "hcx_ref_read_indices": (C.c_int, [C.c_char_p, C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    "hcx_ref_read_cmap": (C.c_int, [C.c_char_p] + [C.POINTER(C.c_int)] * 4 + [C.c_void_p, C.c_void_p, C.c_int]),
    "hcx_ref_ph_to_p2c": (C.c_int, [C.c_int] * 4 + [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.POINTER(C.c_int),
                                   C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "hcx_ref_write_converged": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                          C.c_uint32]),
}


def load(path=LIB_PATH):
    if not os.path.exists(path):
        raise RuntimeError(f"libhcx.so not built at {path}; run __graft_entry__.build() or make -C the package")
    # One HIP runtime per process: torch ships its own libamdhip64 (soname
    # libamdhip64.so.7) and requests it as "libamdhip64.so", so loading libhcx
    # first would map a second runtime.  Loading torch first makes libhcx's
    # libamdhip64.so.7 dependency resolve to the already-mapped runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib
