"""ctypes binding of libzsc.so (declarations mirror include/zsc.h).

The library is built in-tree by ``lbm_bench_gpu_amd/csrc/Makefile``; loading
fails loudly when it is missing -- there is no Python or CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libzsc.so")

ZSC_OK = 0
ZSC_ERR_ARG = -1
ZSC_ERR_IO = -2
ZSC_ERR_HIP = -3
ZSC_ERR_ALLOC = -4
ZSC_ERR_STATE = -5
ZSC_ERR_DIVERGED = -6
ZSC_ERR_PRECISION = -7

PRECISIONS = {"f64": 0, "f32": 1, "bf16": 2, "fp16": 3, "f32d": 4}

_D, _I, _VP = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
_DP = ctypes.POINTER(ctypes.c_double)


class ZscInputs(ctypes.Structure):
    _fields_ = ([(n, _I) for n in ("max_step", "relax_step", "save", "stat", "xmax", "ymax",
                                   "zmax", "block_x", "block_y")] +
                [(n, _D) for n in ("rho_l", "rho_h", "tau_rho", "tau_phi", "width", "sigma",
                                   "gamma", "eo", "phi_star", "gravity")])


PARAM_FIELDS = ("phi_star", "phi_star_sq", "rho_mean", "alpha", "alpha4", "kappa", "gamma",
                "eta", "eta2", "inv_eta2", "inv_tau_phi", "inv_tau_phi_one", "inv_tau_rho_one",
                "k0", "k1", "k2", "kc0", "kc1", "kc2", "gravity", "grav", "width", "sigma")


class ZscParams(ctypes.Structure):
    _fields_ = [(n, _D) for n in PARAM_FIELDS]


STATS_FIELDS = ("sum_phi", "sum_rho", "bubble_volume", "ux_sum", "uy_sum", "uz_sum", "umax_sq",
                "p_in_sum", "p_out_sum", "bubble_volume_full")


class ZscStats(ctypes.Structure):
    _fields_ = [(n, _D) for n in STATS_FIELDS]


_PP7 = _VP * 7
_PP19 = _VP * 19

# name -> (restype, argtypes); every symbol include/zsc.h declares.
SIGNATURES = {
    "zsc_read_inputs": (_I, [ctypes.c_char_p, ctypes.POINTER(ZscInputs)]),
    "zsc_read_bubbles": (_I, [ctypes.c_char_p, _DP, _I, ctypes.POINTER(_I)]),
    "zsc_write_bubbles": (_I, [ctypes.c_char_p, _DP, _I]),
    "zsc_derive_params": (_I, [ctypes.POINTER(ZscInputs), _D, ctypes.POINTER(ZscParams)]),
    "zsc_generate_bubbles": (_I, [ctypes.c_uint64, _I, _D, _D, _D, _D, _I, _I, _I, _I, _DP]),
    "zsc_init_phi_host": (_I, [_I, _I, _I, _DP, _I, _D, _D, _DP]),
    "zsc_create": (_I, [_I, _I, _I, _I, _VP, ctypes.POINTER(_VP)]),
    "zsc_destroy": (_I, [_VP]),
    "zsc_set_params": (_I, [_VP, ctypes.POINTER(ZscParams)]),
    "zsc_set_gravity": (_I, [_VP, _I]),
    "zsc_set_fused_config": (_I, [_VP, _I, _I, _I]),
    "zsc_get_fused_config": (_I, [_VP, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "zsc_sync": (_I, [_VP]),
    "zsc_last_error": (ctypes.c_char_p, []),
    "zsc_init_f": (_I, [_VP, _VP, _PP7]),
    "zsc_init_g": (_I, [_VP, _VP, _VP, _PP19]),
    "zsc_update_phi": (_I, [_VP, _VP, _PP7]),
    "zsc_update_rho": (_I, [_VP, _VP, _PP19]),
    "zsc_update_velocity": (_I, [_VP, _VP, _VP, _VP, _VP, _PP19]),
    "zsc_collision_f": (_I, [_VP, _VP, _VP, _VP, _VP, _VP, _PP7]),
    "zsc_collision_g": (_I, [_VP, _VP, _VP, _VP, _VP, _VP, _PP19]),
    "zsc_stream_f": (_I, [_VP, _PP7]),
    "zsc_stream_g": (_I, [_VP, _PP19]),
    "zsc_state_init": (_I, [_VP, _DP, _DP]),
    "zsc_state_init_bubbles": (_I, [_VP, _DP, _I]),
    "zsc_step": (_I, [_VP, _I]),
    "zsc_step_timed": (_I, [_VP, _I, _DP]),
    "zsc_place_lattices": (_I, [_VP, _I, _DP]),
    "zsc_state_get": (_I, [_VP, _DP, _DP]),
    "zsc_state_set": (_I, [_VP, _DP, _DP]),
    "zsc_state_moments": (_I, [_VP, _DP, _DP, _DP, _DP, _DP]),
    "zsc_state_export": (_I, [_VP, _PP7, _PP19]),
    "zsc_state_import": (_I, [_VP, _PP7, _PP19]),
    "zsc_state_stats": (_I, [_VP, ctypes.POINTER(ZscStats)]),
    "zsc_moment_stats": (_I, [_VP, _VP, _VP, _VP, _VP, _VP, ctypes.POINTER(ZscStats)]),
    "zsc_moment_stats_noalloc": (_I, [_VP, _VP, _VP, _VP, _VP, _VP, ctypes.POINTER(ZscStats)]),
    "zsc_state_pressure": (_I, [_VP, _DP]),
    "zsc_vtk_save": (_I, [_VP, ctypes.c_char_p, _I]),
    "zsc_state_save": (_I, [_VP, ctypes.c_char_p, ctypes.c_int64]),
    "zsc_state_load": (_I, [_VP, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)]),
    "zsc_state_save_ex": (_I, [_VP, ctypes.c_char_p, ctypes.c_int64, _DP, _I]),
    "zsc_state_load_ex": (_I, [_VP, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64), _DP, _I,
                               ctypes.POINTER(_I)]),
    "zsc_bytes_per_cell": (_D, [_I]),
    "zsc_copy_ceiling": (_I, [_VP, _DP]),
    "zsc_adj_scratch_bytes": (ctypes.c_size_t, [_VP]),
    "zsc_adj_set_lambda_precision": (_I, [_VP, _I]),
    "zsc_adj_init": (_I, [_VP, _VP, _VP]),
    "zsc_adj_forward": (_I, [_VP, _VP, _VP, _VP]),
    "zsc_adj_seed": (_I, [_VP, _VP, _VP, _VP, _VP, _VP]),
    "zsc_adj_backward": (_I, [_VP, _VP, _I, _VP, _VP, _VP, _VP, _VP]),
    "zsc_adj_backward_steps": (_I, [_VP, _I, _VP, _VP, _I, _VP, _VP, _VP, _VP, _VP]),
    "zsc_adj_flush": (_I, [_VP, _VP, _VP]),
    "zsc_adj_collide": (_I, [_VP, _VP, _VP, _VP]),
    "zsc_adj_fused": (_I, [_VP, _VP, _VP, _VP]),
    "zsc_adj_stream": (_I, [_VP, _VP, _VP]),
    "zsc_adj_init_backward": (_I, [_VP, _VP, _VP, _VP, _VP]),
}

RESTART_USER_MAX = 8  # ZSC_RESTART_USER_MAX

_lib = None


class ZscError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libzsc error {code}: {msg}")
        self.code = code


def lib():
    """Loads libzsc.so once; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `make -C lbm_bench_gpu_amd/csrc` "
                "(or __graft_entry__.build()); there is no fallback implementation")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(code: int) -> int:
    if code != ZSC_OK:
        raise ZscError(code, lib().zsc_last_error().decode(errors="replace"))
    return code
