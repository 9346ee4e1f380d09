"""Loader of libabie_hip.so with explicit ctypes signatures for every entry
point declared in include/abie_hip.h.

The library is built in-tree (abie_amd/libabie_hip.so, see abie_amd/build.py).
There is no fallback: if it is missing, loading raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libabie_hip.so")

c_dp = ctypes.POINTER(ctypes.c_double)
c_fp = ctypes.POINTER(ctypes.c_float)
c_ip = ctypes.POINTER(ctypes.c_int)
c_llp = ctypes.POINTER(ctypes.c_longlong)
c_vp = ctypes.c_void_p
c_sz = ctypes.c_size_t
c_d = ctypes.c_double
c_i = ctypes.c_int

# name: (restype, argtypes) -- mirrors include/abie_hip.h
SIGNATURES = {
    # reference ABI (src/common.h:74-122)
    "initialize_code": (c_i, [c_d, c_d, c_sz, c_sz, c_sz, c_i]),
    "finalize_code": (c_i, []),
    "set_state": (None, [c_dp, c_dp, c_dp, c_dp, c_sz, c_d, c_d]),
    "get_state": (c_i, [c_dp, c_dp, c_dp, c_dp]),
    "get_model_time": (c_d, []),
    "set_close_encounter_distance": (None, [c_d]),
    "get_close_encounter_distance": (c_d, []),
    "get_close_encounter_buffer": (None, [c_dp]),
    "get_collision_buffer": (None, [c_dp]),
    "reset_close_encounter_buffer": (None, []),
    "reset_collision_buffer": (None, []),
    "set_additional_forces": (c_sz, [c_sz, c_dp]),
    "calculate_energy": (c_d, []),
    "integrator_gr": (c_i, [c_d, c_d, c_d]),
    "integrator_gauss_radau15": (c_sz, [c_dp, c_dp, c_dp, c_dp, c_sz, c_d, c_d, c_d, c_d]),
    "calculate_accelerations": (c_sz, [c_dp, c_dp, c_sz, c_d, c_dp, c_dp, c_dp]),
    "ode_n_body_second_order": (c_sz, [c_dp, c_sz, c_d, c_dp, c_dp, c_dp]),
    "ode_n_body_second_order_gpu": (c_sz, [c_dp, c_sz, c_d, c_dp, c_dp, c_dp]),
    "check_collisions_close_encounters": (c_sz, [c_dp, c_dp, c_sz, c_d]),
    "gpu_init": (None, [c_i, c_i]),
    "gpu_finalize": (None, []),
    "integrator_rk": (c_i, [c_d, c_d, c_d]),
    "integrator_wh": (c_i, [c_d, c_d, c_d]),
    "integrator_runge_kutta": (None, [c_dp, c_dp, c_dp, c_dp, c_sz, c_d, c_d, c_d, c_d]),
    "integrator_wisdom_holman": (None, [c_dp, c_dp, c_dp, c_dp, c_sz, c_d, c_d, c_d, c_d]),
    "ode_n_body_first_order": (c_sz, [c_dp, c_sz, c_d, c_dp, c_dp]),
    # device entry points
    "abie_hip_last_error": (ctypes.c_char_p, []),
    "abie_hip_set_device": (c_i, [c_i]),
    "abie_hip_set_force_mode": (c_i, [c_i]),
    "abie_hip_accel_soa": (c_i, [c_vp, c_vp, c_vp, c_vp, c_i, c_vp, c_vp, c_vp, c_i, c_i, c_vp, c_vp, c_vp, c_i, c_vp]),
    "abie_hip_accel_soa_noalloc": (c_i, [c_vp, c_vp, c_vp, c_vp, c_i, c_vp, c_vp, c_vp, c_i, c_i, c_vp, c_vp, c_vp,
                                         c_i, c_vp]),
    "abie_hip_issue_rate": (c_i, [c_i, ctypes.c_double, c_vp, c_vp]),
    "abie_hip_step_root7": (c_d, [c_d]),
    "abie_hip_time_accel_soa": (c_i, [c_vp, c_vp, c_vp, c_vp, c_i, c_vp, c_vp, c_vp, c_i, c_i, c_vp, c_vp, c_vp,
                                      c_i, c_vp, c_i, c_dp, c_dp, c_dp, c_dp]),
    "abie_ensemble_create": (c_i, [ctypes.POINTER(c_vp), c_i, c_ip, c_dp, c_dp, c_dp, c_dp, c_d, c_i, c_i]),
    "abie_ensemble_destroy": (None, [c_vp]),
    "abie_ensemble_run": (c_i, [c_vp, c_d, c_d, c_ip]),
    "abie_ensemble_get_state": (c_i, [c_vp, c_dp, c_dp, c_dp, c_ip]),
    "abie_ensemble_get_counters": (c_i, [c_vp, c_llp]),
    "abie_ensemble_energy": (c_i, [c_vp, c_dp]),
    "abie_ensemble_get_finish_times": (c_i, [c_vp, c_dp]),
    "abie_ensemble_set_steplog": (c_i, [c_vp, ctypes.c_longlong]),
    "abie_ensemble_get_steplog": (c_i, [c_vp, c_dp, c_llp]),
    "abie_ensemble_set_proposer": (c_i, [c_vp, c_fp, c_i]),
    "abie_gr15_num_configs": (c_i, []),
    "abie_gr15_config": (c_i, [c_i, c_ip, c_ip]),
    "abie_gr15_set_tuning": (c_i, [c_i, c_ip, c_ip]),
    "abie_mlp_forward": (c_i, [c_vp, c_vp, c_vp, c_i, c_vp]),
    "abie_mlp_loss_grad": (c_i, [c_vp, c_vp, c_vp, c_vp, c_i, c_vp, c_vp, c_vp]),
    "abie_mlp_adam": (c_i, [c_vp, c_vp, c_vp, c_vp, c_i, c_i, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                            ctypes.c_float, c_vp]),
}

# constants of include/abie_hip.h
FORCE_FP64, FORCE_FP64_EXACT, FORCE_MIXED, FORCE_FP32 = 0, 1, 2, 3
CTRL_PLAIN, CTRL_MASSIVE, CTRL_LEARNED = 0, 1, 2
STEPLOG_FIELDS = 31
MLP_IN, MLP_HIDDEN = 12, 64
MLP_NPARAM = MLP_IN * MLP_HIDDEN + MLP_HIDDEN + MLP_HIDDEN * MLP_HIDDEN + MLP_HIDDEN + MLP_HIDDEN + 1

_lib = None


class AbieHipError(RuntimeError):
    pass


def load():
    """Load libabie_hip.so (raises if it has not been built).

    When PyTorch is importable it is imported first: torch ships its own HIP
    runtime, and loading it before this library makes both use that one
    runtime (same soname) instead of two in one process."""
    global _lib
    if _lib is None:
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise AbieHipError("libabie_hip.so not built: run `python -m abie_amd.build` (%s)" % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        _apply_tuning(lib)
    return _lib


TUNING_PATH = os.path.join(HERE, "data", "gr15_tuning.json")


def _apply_tuning(lib):
    """Use the autotuner's rules table (abie_amd/autotune.py) when present."""
    import json
    try:
        with open(os.environ.get("ABIE_HIP_TUNING", TUNING_PATH)) as f:
            rules = json.load(f)["rules"]
    except (OSError, ValueError, KeyError):
        return
    mx = (ctypes.c_int * len(rules))(*[int(r[0]) for r in rules])
    cf = (ctypes.c_int * len(rules))(*[int(r[1]) for r in rules])
    lib.abie_gr15_set_tuning(len(rules), mx, cf)


def check(rc, what="libabie_hip"):
    if rc is not None and int(rc) < 0:
        msg = load().abie_hip_last_error()
        raise AbieHipError("%s failed (%d): %s" % (what, int(rc), msg.decode() if msg else ""))
    return rc
