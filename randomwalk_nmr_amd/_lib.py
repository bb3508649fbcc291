"""ctypes binding of librwnmr_amd.so (the C ABI in include/rwnmr/*.h).

There is no fallback: if the library is missing the import of any compute
entry point raises, so a GPU run can never silently use another path.
"""
from __future__ import annotations

import ctypes as ct
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librwnmr_amd.so")

P = ct.c_void_p
I32 = ct.c_int32
U32 = ct.c_uint32
U64 = ct.c_uint64
I64 = ct.c_int64
D = ct.c_double
CI = ct.c_int
B = ct.c_bool

RWNMR_OK = 0
STATUS_NAMES = {0: "OK", 1: "INVALID", 2: "HIP", 3: "ALLOC", 4: "STATE", 5: "WALKER", 6: "NUMERIC"}


class RwnmrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rwnmr error {STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


class Walkers(ct.Structure):
    """rwnmr_walkers (include/rwnmr/rwnmr.h)."""
    _fields_ = [("x", P), ("y", P), ("z", P), ("seed", P), ("energy", P), ("hits", P),
                ("penalty", P), ("delta_uniform", D), ("n", U64), ("on_device", I32)]


class UctConfig(ct.Structure):
    """rwnmr_uct_config (include/rwnmr/rwnmr_io.h)."""
    _fields_ = [("dir_path", ct.c_char * 1024), ("filename", ct.c_char * 256), ("first_idx", I32),
                ("digits", I32), ("extension", ct.c_char * 32), ("slices", I32), ("resolution", D),
                ("voxel_division", I32), ("img_files_list", ct.c_char * 1024)]


class CpmgInstance(ct.Structure):
    """rwnmr_cpmg_instance (include/rwnmr/rwnmr.h)."""
    _fields_ = [("image", P), ("walkers", Walkers), ("echoes", U32), ("steps_per_echo", U32),
                ("bc", I32), ("echo_sums", P)]


class CpmgParams(ct.Structure):
    """rwnmr_cpmg_params (include/rwnmr/rwnmr_drivers.h)."""
    _fields_ = [("observation_time", D), ("steps_per_echo", U32), ("voxel_size", D), ("D0", D),
                ("bulk_time", D), ("apply_bulk", I32), ("rho", D), ("sigmoid", P), ("bc", I32), ("shift", U32)]


class PfgseParams(ct.Structure):
    """rwnmr_pfgse_params (include/rwnmr/rwnmr_drivers.h)."""
    _fields_ = [("voxel_size", D), ("D0", D), ("bulk_time", D), ("apply_bulk", I32), ("steps_per_echo", U32),
                ("bc", I32), ("pulse_width", D), ("rho", D), ("sigmoid", P), ("n_rho_sets", U32),
                ("rho_sets", P), ("k_per_delta", I32), ("threshold_samples", U32), ("threshold_lhs", D),
                ("threshold_window", U32)]


class MultitauParams(ct.Structure):
    """rwnmr_multitau_params (include/rwnmr/rwnmr_drivers.h)."""
    _fields_ = [("tau_min", D), ("tau_max", D), ("tau_points", U32), ("tau_scale", I32), ("tau_values", P),
                ("complete_decay", I32)]


class PfgseResult(ct.Structure):
    """rwnmr_pfgse_result (include/rwnmr/rwnmr_drivers.h)."""
    _fields_ = [("Mkt", P), ("rhs", P), ("lhs", P), ("D_sat", P), ("D_sat_err", P), ("points", P), ("msd", P),
                ("D_msd", P), ("steps", P)]


# (name, restype, argtypes) of every exported symbol the package binds.
SIGNATURES = {
    # image ingestion and text outputs (include/rwnmr/rwnmr_io.h)
    "rwnmr_uct_config_defaults": (None, [P]),
    "rwnmr_uct_config_read": (I32, [ct.c_char_p, P]),
    "rwnmr_uct_slice_path": (I32, [P, I32, ct.c_char_p, U32]),
    "rwnmr_uct_dims": (I32, [P, ct.POINTER(I32), ct.POINTER(I32), ct.POINTER(I32)]),
    "rwnmr_uct_load": (I32, [P, I32, I32, I32, P]),
    "rwnmr_png_size": (I32, [ct.c_char_p, ct.POINTER(I32), ct.POINTER(I32)]),
    "rwnmr_png_binarize": (I32, [ct.c_char_p, I32, I32, P]),
    "rwnmr_image_stats": (I32, [P, I32, I32, I32, D, ct.POINTER(U64), ct.POINTER(D), ct.POINTER(U64),
                                ct.POINTER(D)]),
    "rwnmr_write_image_info": (I32, [ct.c_char_p, ct.c_char_p, I32, I32, I32, D, I32, D, D, D]),
    "rwnmr_write_walkers_csv": (I32, [ct.c_char_p, U64, P, P, P, P, P, P, P, P, P, P]),
    "rwnmr_version": (ct.c_char_p, []),
    "rwnmr_ctx_create": (I32, [ct.POINTER(P), I32]),
    "rwnmr_ctx_destroy": (None, [P]),
    "rwnmr_last_error": (ct.c_char_p, [P]),
    "rwnmr_ctx_synchronize": (I32, [P]),
    "rwnmr_ctx_stream": (P, [P]),
    "rwnmr_ctx_set_launch": (I32, [P, U32, U32]),
    "rwnmr_ctx_get_launch": (I32, [P, ct.POINTER(U32), ct.POINTER(U32)]),
    "rwnmr_ctx_set_image_layout": (I32, [P, I32, I32, I32]),
    "rwnmr_ctx_set_locality": (I32, [P, I32, U32, U64]),
    "rwnmr_ctx_set_unfolded": (I32, [P, I32]),
    "rwnmr_ctx_set_timing": (I32, [P, I32]),
    "rwnmr_ctx_get_timing": (I32, [P, ct.POINTER(D), ct.POINTER(U64), I32]),
    "rwnmr_ctx_autotune": (I32, [P, P, ct.POINTER(Walkers), U32, I32, U32, ct.POINTER(U32), ct.POINTER(U32),
                                 ct.POINTER(D)]),
    "rwnmr_ctx_get_tuning": (I32, [P, ct.POINTER(U32), ct.POINTER(U32), ct.POINTER(I32), ct.POINTER(U32)]),
    "rwnmr_image_upload": (I32, [P, P, I32, I32, I32, P, I32, ct.POINTER(P)]),
    "rwnmr_image_destroy": (None, [P]),
    "rwnmr_bitblock_pack": (I32, [P, I32, I32, I32, P]),
    "rwnmr_map_run": (I32, [P, P, ct.POINTER(Walkers), U64, I32, U32]),
    "rwnmr_cpmg_run": (I32, [P, P, ct.POINTER(Walkers), U32, U32, I32, U32, P]),
    "rwnmr_cpmg_run_noalloc": (I32, [P, P, ct.POINTER(Walkers), U32, U32, I32, U32, P]),
    "rwnmr_cpmg_graph_create": (I32, [P, P, ct.POINTER(Walkers), ct.POINTER(Walkers), U32, U32, I32, P,
                                      ct.POINTER(P)]),
    "rwnmr_graph_launch": (I32, [P]),
    "rwnmr_graph_destroy": (None, [P]),
    "rwnmr_cpmg_run_labels": (I32, [P, P, ct.POINTER(Walkers), U32, U32, I32, U32, P, P, P]),
    "rwnmr_cpmg_run_label_sets": (I32, [P, P, ct.POINTER(Walkers), U32, U32, I32, U32, P, P]),
    "rwnmr_pfgse_run_sampled": (I32, [P, P, ct.POINTER(Walkers), P, P, P, P, U32, P, P, P, U32, I32, D, I32, U32, P,
                                      U32, P, P, P]),
    "rwnmr_pfgse_sample_stats": (I32, [P, U32, U32, P, U32, P, P, P, P, P, ct.POINTER(D), ct.POINTER(D),
                                       ct.POINTER(D)]),
    "rwnmr_multitau_sequence": (I32, [ct.POINTER(MultitauParams), D, P, P]),
    "rwnmr_multitau_simulate": (I32, [P, P, ct.POINTER(Walkers), ct.POINTER(MultitauParams),
                                      ct.POINTER(CpmgParams), P, P]),
    "rwnmr_config_create": (I32, [ct.POINTER(P)]),
    "rwnmr_config_destroy": (None, [P]),
    "rwnmr_config_read": (I32, [P, ct.c_char_p]),
    "rwnmr_config_set": (I32, [P, ct.c_char_p, ct.c_char_p]),
    "rwnmr_config_has": (I32, [P, ct.c_char_p]),
    "rwnmr_config_string": (I32, [P, ct.c_char_p, ct.c_char_p, U32]),
    "rwnmr_config_double": (I32, [P, ct.c_char_p, ct.POINTER(D)]),
    "rwnmr_config_int": (I32, [P, ct.c_char_p, ct.POINTER(I64)]),
    "rwnmr_config_bool": (I32, [P, ct.c_char_p, ct.POINTER(I32)]),
    "rwnmr_config_list": (I32, [P, ct.c_char_p, I32, P, U32, ct.POINTER(U32)]),
    "rwnmr_histogram_bins": (I32, [I32, I32]),
    "rwnmr_histogram_list": (I32, [U32, I32, ct.POINTER(I32), P, P]),
    "rwnmr_collision_histogram": (I32, [P, U64, U64, I32, I32, P, P]),
    "rwnmr_histogram_penalties": (I32, [P, I32, D, P, D, D, U32, P]),
    "rwnmr_histogram_decay": (I32, [P, I32, I32, P, P, P, P]),
    "rwnmr_cpmg_histogram_simulate": (I32, [P, P, ct.POINTER(Walkers), U32, U32, I32, U32, I32, I32, I32, D, P, D,
                                            D, P, P, P, ct.POINTER(I32)]),
    "rwnmr_cpmg_field_run": (I32, [P, P, ct.POINTER(Walkers), U32, U32, I32, P, D, D, P, P, P]),
    "rwnmr_internal_field_uniform": (I32, [P, I32, I32, I32, D, D, I32, P]),
    "rwnmr_internal_field_read": (I32, [ct.c_char_p, I32, I32, I32, P]),
    "rwnmr_pfgse_run": (I32, [P, P, ct.POINTER(Walkers), P, P, P, P, U32, P, P, P, U32, I32, D, I32, U32,
                              U32, P, P, P]),
    "rwnmr_cpmg_batch_run": (I32, [P, ct.POINTER(CpmgInstance), U32, U32]),
    "rwnmr_cpmg_forward_ckpt": (I32, [P, P, ct.POINTER(Walkers), U32, U32, I32, P, U32, P, P]),
    "rwnmr_cpmg_forward_jacobian": (I32, [P, P, ct.POINTER(Walkers), U32, U32, I32, P, P, P, P]),
    "rwnmr_cpmg_backward": (I32, [P, P, P, U64, U32, U32, I32, P, U32, P, P]),
    "rwnmr_ckpt_bytes": (U64, [U64, U32, U32]),
    "rwnmr_probe_gather": (I32, [P, U64, U32, ct.POINTER(D), ct.POINTER(D)]),
    # sequence drivers (include/rwnmr/rwnmr_drivers.h)
    "rwnmr_cpmg_simulate": (I32, [P, P, ct.POINTER(Walkers), ct.POINTER(CpmgParams), U32, P, P,
                                  ct.POINTER(U32)]),
    "rwnmr_pfgse_simulate": (I32, [P, P, ct.POINTER(Walkers), ct.POINTER(PfgseParams), P, U32, P, P, P, U32,
                                   ct.POINTER(PfgseResult)]),
    "rwnmr_pfgse_gradients": (None, [D, D, D, U32, P, P, P]),
    "rwnmr_pfgse_wave_vector": (D, [D, D, D]),
    "rwnmr_stejskal_tanner": (I32, [P, P, U32, ct.POINTER(D), ct.POINTER(D)]),
    "rwnmr_pfgse_threshold_lhs": (U32, [P, U32, D, U32]),
    "rwnmr_pfgse_threshold_lhs_window": (U32, [P, U32, D, U32]),
    # output tree (include/rwnmr/rwnmr_output.h)
    "rwnmr_make_dirs": (I32, [ct.c_char_p]),
    "rwnmr_write_sim_info": (I32, [ct.c_char_p, ct.c_char_p, ct.c_char_p, ct.c_char_p, D, D, I32, D, U64, I32,
                                   ct.c_char_p]),
    "rwnmr_write_bin_image": (I32, [ct.c_char_p, P, I32, I32, I32]),
    "rwnmr_write_cpmg_decay": (I32, [ct.c_char_p, U32, P, P, P]),
    "rwnmr_write_two_columns": (I32, [ct.c_char_p, ct.c_char_p, U32, P, P]),
    "rwnmr_write_histogram_list": (I32, [ct.c_char_p, I32, I32, P, P]),
    "rwnmr_write_pfgse_parameters": (I32, [ct.c_char_p, D, D, D, D, D, D, U32, P, U32]),
    "rwnmr_write_pfgse_gradient": (I32, [ct.c_char_p, U32, P, P, P, P, P, P]),
    "rwnmr_write_pfgse_echoes": (I32, [ct.c_char_p, U32, P, P, P, P, P, P, P]),
    "rwnmr_write_pfgse_msd": (I32, [ct.c_char_p, P, P, P, P]),
    "rwnmr_write_pfgse_results_header": (I32, [ct.c_char_p]),
    "rwnmr_append_pfgse_results": (I32, [ct.c_char_p, D, P, U32]),
    # host set-up (include/rwnmr/rwnmr_host.h)
    "rwnmr_synth_sphere_pack": (I32, [I32, I32, I32, D, D, D, U64, P, ct.POINTER(D)]),
    "rwnmr_synth_voronoi_labels": (I32, [I32, I32, I32, I32, I32, U64, P, P]),
    "rwnmr_count_pores": (I64, [P, I32, I32, I32]),
    "rwnmr_seed_and_place": (I32, [P, I32, I32, I32, U64, U64, P, P, P, P]),
    "rwnmr_setup_walkers": (I32, [P, I32, I32, I32, U64, U64, I32, I32, U32, P, P, P, P, ct.POINTER(D),
                                  ct.POINTER(U64)]),
    "rwnmr_time_step": (D, [D, D]),
    "rwnmr_time_framework": (I32, [D, D, D, U32, ct.POINTER(U32), ct.POINTER(U64)]),
    "rwnmr_penalty": (D, [D, D, D]),
    # Laplace inversion, reference C API names (B/Laplace/include/nmrinv.h)
    "initialize_inverter": (None, [CI, P, D, D, CI, CI, CI, D]),
    "initialize_inverter_with_t2_bins": (None, [CI, P, CI, P, CI, CI, D]),
    "find_best_lambda": (D, [CI, P]),
    "get_processed_raw_amps": (None, [CI, P, P]),
    "invert": (B, [CI, P, P]),
    "invert_with_multiple_noises": (B, [CI, P, CI, P, P]),
    "get_lcurve_values": (B, [P, P, P, P]),
    "get_tikhonov_lambda": (D, []),
    "set_lambda": (None, [D]),
    "get_noise": (B, [P]),
    "set_noise": (None, [D, P]),
    "get_used_raw_bins": (None, [P]),
    "get_used_t2_bins": (None, [P]),
    "rwnmr_inverter_num_used_raw_bins": (CI, []),
    "rwnmr_inverter_num_t2_bins": (CI, []),
    "rwnmr_inverter_num_lambdas": (CI, []),
    "rwnmr_inverter_set_lambda_range": (None, [D, D]),
    "rwnmr_laplace_nnls": (I32, [CI, CI, P, P, P, D, D, CI]),
}

_lib = None


def load():
    """Load librwnmr_amd.so and bind every signature; raises if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m randomwalk_nmr_amd.build` "
                          "(there is no CPU fallback)")
    L = ct.CDLL(LIB_PATH)
    missing = []
    for name, (res, args) in SIGNATURES.items():
        try:
            f = getattr(L, name)
        except AttributeError:
            missing.append(name)
            continue
        f.restype = res
        f.argtypes = args
    L._rwnmr_missing = missing
    _lib = L
    return L


def check(code, ctx=None):
    if code != RWNMR_OK:
        msg = ""
        if ctx is not None:
            m = load().rwnmr_last_error(ctx)
            msg = m.decode() if m else ""
        raise RwnmrError(code, msg)
