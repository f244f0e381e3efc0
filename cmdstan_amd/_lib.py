"""ctypes binding of libcmdstan_amd.so (the C ABI declared in include/cmdstan_amd.h).

The library is built in-tree (cmdstan_amd/libcmdstan_amd.so).  Loading fails
loudly when it is missing: there is no CPU fallback for the sampler.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcmdstan_amd.so")

DP = C.POINTER(C.c_double)
I64P = C.POINTER(C.c_int64)


INTERRUPT_FN = C.CFUNCTYPE(C.c_int, C.c_void_p)


class nuts_config(C.Structure):
    _fields_ = [
        ("num_samples", C.c_int), ("num_warmup", C.c_int), ("save_warmup", C.c_int), ("thin", C.c_int),
        ("adapt_engaged", C.c_int), ("gamma", C.c_double), ("delta", C.c_double), ("kappa", C.c_double),
        ("t0", C.c_double), ("init_buffer", C.c_int), ("term_buffer", C.c_int), ("window", C.c_int),
        ("max_depth", C.c_int), ("metric", C.c_int), ("stepsize", C.c_double), ("stepsize_jitter", C.c_double),
        ("num_chains", C.c_int), ("id", C.c_int), ("seed", C.c_uint64), ("init_radius", C.c_double),
        ("fused", C.c_int), ("precision", C.c_int), ("max_saved_params", C.c_int),
        ("flow", C.c_int), ("flow_layers", C.c_int), ("flow_hidden", C.c_int), ("flow_train_steps", C.c_int),
        ("flow_batch", C.c_int), ("flow_lr", C.c_double), ("flow_warmup", C.c_int), ("flow_pilot", C.c_int),
        ("refresh", C.c_int), ("interrupt", INTERRUPT_FN), ("interrupt_user", C.c_void_p),
    ]


class nuts_output(C.Structure):
    _fields_ = [
        ("draws", DP), ("stepsize", DP), ("inv_metric", DP), ("n_leapfrog", I64P), ("grad_evals", I64P),
        ("status", C.POINTER(C.c_int)), ("warmup_seconds", C.c_double), ("sampling_seconds", C.c_double),
        ("leaf_steps", C.c_int64), ("grad_seconds_per_launch", C.c_double), ("grad_launches_timed", C.c_int64),
        ("grad_chains_per_launch", C.c_double),
        ("flow_used", C.c_int), ("flow_train_seconds", C.c_double), ("flow_loss", C.c_double),
        ("pilot_ess_per_grad_plain", C.c_double), ("pilot_ess_per_grad_flow", C.c_double),
        ("pilot_ess_per_sec_plain", C.c_double), ("pilot_ess_per_sec_flow", C.c_double),
        ("pilot_rhat_plain", C.c_double), ("pilot_rhat_flow", C.c_double),
        ("tuned_path", C.c_int), ("tuned_blocks_per_cu", C.c_int), ("diagnostics", DP),
        ("energy_error", DP), ("init_rejections", C.POINTER(C.c_int)),
        ("tuned_threads_per_chain", C.c_int), ("tuned_wave_launch_seconds", C.c_double),
        ("tuned_block_launch_seconds", C.c_double),
    ]


PLUGIN_GRAD = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                          C.c_void_p, C.c_void_p, C.c_void_p)
PLUGIN_WRITE = C.CFUNCTYPE(C.c_int, C.c_void_p, DP, C.c_int, DP)
PLUGIN_INITS = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, DP)


class nuts_model_plugin(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("num_outputs", C.c_int), ("param_names", C.POINTER(C.c_char_p)),
        ("unconstrained_names", C.POINTER(C.c_char_p)), ("user", C.c_void_p),
        ("log_prob_grad", PLUGIN_GRAD), ("write_array", PLUGIN_WRITE), ("transform_inits", PLUGIN_INITS),
    ]


# every entry point the header declares, with its ctypes signature
SIGNATURES = {
    "nuts_data_create": (C.c_void_p, []),
    "nuts_data_destroy": (None, [C.c_void_p]),
    "nuts_data_add": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, I64P, DP, C.c_int]),
    "nuts_data_parse_json": (C.c_int, [C.c_void_p, C.c_char_p]),
    "nuts_data_parse_rdump": (C.c_int, [C.c_void_p, C.c_char_p]),
    "nuts_data_get": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_int), I64P, C.POINTER(DP)]),
    "nuts_model_create": (C.c_void_p, [C.c_char_p, C.POINTER(C.c_void_p), C.c_int]),
    "nuts_model_create_plugin": (C.c_void_p, [C.POINTER(nuts_model_plugin)]),
    "nuts_model_destroy": (None, [C.c_void_p]),
    "nuts_model_num_datasets": (C.c_int, [C.c_void_p]),
    "nuts_model_dims": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "nuts_model_param_name": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_char_p, C.c_int]),
    "nuts_model_unconstrained_name": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_char_p, C.c_int]),
    "nuts_log_prob_grad": (C.c_int, [C.c_void_p, C.c_int, DP, C.c_int, C.c_int, DP, DP]),
    "nuts_log_prob_grad_precision": (C.c_int, [C.c_void_p, C.c_int, DP, C.c_int, C.c_int, C.c_int, DP, DP]),
    "nuts_log_prob_grad_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                             C.c_void_p, C.c_void_p]),
    "nuts_model_transform_inits": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, DP]),
    "nuts_config_default": (None, [C.POINTER(nuts_config)]),
    "nuts_config_validate": (C.c_int, [C.POINTER(nuts_config)]),
    "nuts_output_rows": (C.c_int, [C.POINTER(nuts_config)]),
    "nuts_sample": (C.c_int, [C.c_void_p, C.POINTER(nuts_config), DP, DP, C.POINTER(nuts_output)]),
    "nuts_transition_replay": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, DP,
                                         DP, DP, C.c_int64, DP, C.c_int64, DP, DP]),
    "nuts_transition_replay_precision": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                                   C.c_int, DP, DP, DP, C.c_int64, DP, C.c_int64, C.c_int, DP, DP]),
    "nuts_transition_replay_driver": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                                C.c_int, DP, DP, DP, C.c_int64, DP, C.c_int64, C.c_int, C.c_int, DP,
                                                DP]),
    "nuts_summary": (C.c_int, [DP, C.c_int, C.c_int, C.c_int, C.c_int, DP]),
    "nuts_flow_create": (C.c_void_p, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64]),
    "nuts_flow_destroy": (None, [C.c_void_p]),
    "nuts_flow_num_params": (C.c_int64, [C.c_void_p]),
    "nuts_flow_get_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "nuts_flow_set_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "nuts_flow_set_standardization": (C.c_int, [C.c_void_p, DP, DP]),
    "nuts_flow_forward": (C.c_int, [C.c_void_p, DP, C.c_int, DP, DP]),
    "nuts_flow_inverse": (C.c_int, [C.c_void_p, DP, C.c_int, DP, DP]),
    "nuts_flow_pullback_grad": (C.c_int, [C.c_void_p, DP, DP, C.c_int, DP]),
    "nuts_flow_loss_grad": (C.c_int, [C.c_void_p, DP, C.c_int, DP, C.POINTER(C.c_float)]),
    "nuts_flow_train": (C.c_int, [C.c_void_p, DP, C.c_int64, C.c_int, C.c_int, C.c_double, C.c_uint64, DP]),
    "nuts_last_error": (C.c_char_p, []),
    "nuts_version": (C.c_char_p, []),
}

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; "
                               f"g.build()'); the sampler has no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def last_error():
    return lib().nuts_last_error().decode()
