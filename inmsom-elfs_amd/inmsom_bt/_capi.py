"""ctypes binding of libinmsom_bt.so (the C ABI in include/inmsom_bt.h).

The library is built in-tree by ``make -C inmsom-elfs_amd/csrc``; loading
fails loudly when it is missing.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libinmsom_bt.so")

OK, EINVAL, EDEVICE, ERANGE, ENOMEM, ESTATE = 0, -1, -2, -3, -4, -5
FP64_EXACT, FP64, FP32 = 0, 1, 2
MODES = {"fp64_exact": FP64_EXACT, "fp64": FP64, "fp32": FP32}

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_u8p = C.POINTER(C.c_uint8)


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("periodicity_x", C.c_int),
                ("periodicity_y", C.c_int), ("full_free_surface", C.c_int),
                ("free_fall_acc", C.c_double), ("ref_den", C.c_double),
                ("time_smooth", C.c_double), ("lbasins", _ip),
                ("lu", _fp), ("luu", _fp), ("luh", _fp), ("lcu", _fp), ("lcv", _fp),
                ("llu", _fp), ("llv", _fp),
                ("dx", _dp), ("dy", _dp), ("dxt", _dp), ("dyt", _dp), ("dxh", _dp),
                ("dyh", _dp), ("dxb", _dp), ("dyb", _dp), ("rlh_s", _dp), ("hhq_rest", _dp)]


class Forcing(C.Structure):
    _fields_ = [("taux", _dp), ("tauy", _dp), ("wflux", _dp), ("mu", _dp), ("mu4", _dp),
                ("rdis", _dp), ("cb_nl", C.c_double), ("ebottom", C.c_double)]


# every symbol the header declares, with its argument types
SIGNATURES = {
    "inmsom_bt_last_error": ([], C.c_char_p),
    "inmsom_bt_gridcon": ([C.c_int, C.c_int, C.c_int, C.c_int, _ip] + [_fp] * 7, C.c_int),
    "inmsom_bt_rayleigh": ([C.POINTER(Grid), C.c_double, _dp], C.c_int),
    "inmsom_bt_init": ([C.POINTER(Grid), C.c_int, C.POINTER(C.c_void_p)], C.c_int),
    "inmsom_bt_finalize": ([C.c_void_p], C.c_int),
    "inmsom_bt_barotropic_dynamics": ([C.c_void_p, C.c_double, C.c_int, C.c_int] + [_dp] * 16,
                                      C.c_int),
    "inmsom_bt_barotropic_dynamics_dev": ([C.c_void_p, C.c_double, C.c_int, C.c_int] + [_dp] * 16,
                                          C.c_int),
    "inmsom_bt_barotropic_dynamics_noalloc": ([C.c_void_p, C.c_double, C.c_int] + [_dp] * 15, C.c_int),
    "inmsom_bt_set_dropin_path": ([C.c_void_p, C.c_int], C.c_int),
    "inmsom_bt_get_dropin_path": ([C.c_void_p, C.c_int, C.POINTER(C.c_int)], C.c_int),
    "inmsom_bt_dropin_timed": ([C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int, _dp], C.c_int),
    "inmsom_bt_get_depths": ([C.c_void_p, C.POINTER(_dp)], C.c_int),
    "inmsom_bt_set_forcing": ([C.c_void_p, C.POINTER(Forcing)], C.c_int),
    "inmsom_bt_set_state": ([C.c_void_p] + [_dp] * 6, C.c_int),
    "inmsom_bt_get_state": ([C.c_void_p] + [_dp] * 6, C.c_int),
    "inmsom_bt_get_means": ([C.c_void_p] + [_dp] * 4, C.c_int),
    "inmsom_bt_model_steps": ([C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int], C.c_int),
    "inmsom_bt_run": ([C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int], C.c_int),
    "inmsom_bt_run_timed": ([C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int, _dp], C.c_int),
    "inmsom_bt_set_launch": ([C.c_void_p, C.c_int], C.c_int),
    "inmsom_bt_set_pipeline": ([C.c_void_p, C.c_int], C.c_int),
    "inmsom_bt_get_pipeline": ([C.c_void_p, C.c_int, C.POINTER(C.c_int)], C.c_int),
    "inmsom_bt_get_launch": ([C.c_void_p, C.POINTER(C.c_int)], C.c_int),
    "inmsom_bt_set_substeps_per_launch": ([C.c_void_p, C.c_int], C.c_int),
    "inmsom_bt_get_substeps_per_launch": ([C.c_void_p, C.c_int, C.POINTER(C.c_int)], C.c_int),
    "inmsom_bt_range_check": ([C.c_void_p, C.POINTER(C.c_longlong)], C.c_int),
    "inmsom_bt_volume": ([C.c_void_p, _dp], C.c_int),
    "inmsom_bt_synchronize": ([C.c_void_p], C.c_int),
    "inmsom_bt_coarsen_state": ([C.c_void_p, C.c_int, _dp, _dp, _dp], C.c_int),
    "inmsom_bt_remove_barotropic": ([C.c_void_p, C.c_int, _dp, _dp, _dp, _dp], C.c_int),
    "inmsom_bt_internal_update": ([C.c_void_p, C.c_double, C.c_int] + [_dp] * 12 + [C.POINTER(_dp), _dp, _dp],
                                  C.c_int),
    "inmsom_bt_remove_barotropic_dev": ([C.c_void_p, C.c_int, _dp, _dp, _dp, _dp], C.c_int),
    "inmsom_bt_internal_update_dev": ([C.c_void_p, C.c_double, C.c_int] + [_dp] * 12 + [C.POINTER(_dp), _dp, _dp],
                                      C.c_int),
    "inmsom_bt_coupling_timed": ([C.c_void_p, C.c_double, C.c_int, C.c_int, _dp], C.c_int),
    "inmsom_bt_forcing_3d": ([C.c_void_p, C.c_int] + [_dp] * 22, C.c_int),
    "inmsom_bt_forcing_3d_dev": ([C.c_void_p, C.c_int] + [_dp] * 22, C.c_int),
    "inmsom_bt_forcing_3d_timed": ([C.c_void_p, C.c_int, C.c_int, _dp], C.c_int),
    "inmsom_bt_wdstd": ([C.c_char_p, C.c_char_p, C.c_int, _fp, _fp] + [C.c_int] * 9, C.c_int),
    "inmsom_bt_rdstd": ([C.c_char_p, C.c_char_p, C.c_int, _fp, _fp] + [C.c_int] * 9, C.c_int),
    "inmsom_bt_wdstd8": ([C.c_char_p, C.c_char_p, C.c_int, _dp, _fp] + [C.c_int] * 9, C.c_int),
    "inmsom_bt_rdstd8": ([C.c_char_p, C.c_char_p, C.c_int, _dp, _fp] + [C.c_int] * 9, C.c_int),
    "inmsom_bt_ctl_write": ([C.c_char_p, C.c_float] + [C.c_int] * 5 + [_dp, C.c_double, C.c_int, _dp, C.c_double,
                                                                      C.c_int, _dp, C.c_double] + [C.c_int] * 6
                            + [C.c_float, C.c_char_p, C.c_char_p], C.c_int),
    "inmsom_bt_write_snapshot": ([C.c_void_p, C.c_char_p, C.c_int] + [C.c_double] * 4 + [C.c_int] * 6
                                 + [C.c_float, _dp], C.c_int),
    "inmsom_bt_baroclinic_dynamics": ([C.c_void_p, C.c_double, C.c_int] + [_dp] * 26, C.c_int),
    "inmsom_bt_baroclinic_dynamics_dev": ([C.c_void_p, C.c_double, C.c_int] + [_dp] * 26, C.c_int),
    "inmsom_bt_baroclinic_timed": ([C.c_void_p, C.c_double, C.c_int, C.c_int, _dp], C.c_int),
    "inmsom_bt_vertical_velocity": ([C.c_void_p, C.c_int] + [_dp] * 7, C.c_int),
    "inmsom_bt_vertical_velocity_dev": ([C.c_void_p, C.c_int] + [_dp] * 7, C.c_int),
    "inmsom_bt_vertical_velocity_timed": ([C.c_void_p, C.c_int, C.c_int, _dp], C.c_int),
    "inmsom_bt_pressure_gradients": ([C.c_void_p, C.c_int] + [_dp] * 7, C.c_int),
    "inmsom_bt_pressure_gradients_dev": ([C.c_void_p, C.c_int] + [_dp] * 7, C.c_int),
    "inmsom_bt_pressure_gradients_timed": ([C.c_void_p, C.c_int, C.c_int, _dp], C.c_int),
    "inmsom_bt_density": ([C.c_void_p, C.c_int] + [_dp] * 6, C.c_int),
    "inmsom_bt_density_dev": ([C.c_void_p, C.c_int] + [_dp] * 6, C.c_int),
    "inmsom_bt_density_timed": ([C.c_void_p, C.c_int, C.c_int, _dp], C.c_int),
    "inmsom_bt_save_state": ([C.c_void_p], C.c_int),
    "inmsom_bt_cnn_set_fused": ([C.c_void_p, C.c_int], C.c_int),
    "inmsom_bt_restore_state": ([C.c_void_p], C.c_int),
    "inmsom_bt_smagorinsky": ([C.c_void_p, C.c_int, C.c_double, C.c_double] + [_dp] * 6, C.c_int),
    "inmsom_bt_smagorinsky_dev": ([C.c_void_p, C.c_int, C.c_double, C.c_double] + [_dp] * 6, C.c_int),
    "inmsom_bt_smagorinsky_timed": ([C.c_void_p, C.c_int, C.c_int, _dp], C.c_int),
    "inmsom_bt_stress_components": ([C.c_void_p, C.c_int] + [_dp] * 4, C.c_int),
    "inmsom_bt_stress_components_dev": ([C.c_void_p, C.c_int] + [_dp] * 4, C.c_int),
    "inmsom_bt_stress_components_timed": ([C.c_void_p, C.c_int, C.c_int, _dp], C.c_int),
    "inmsom_bt_ens_create": ([C.c_int, C.POINTER(Grid), C.POINTER(Forcing), C.c_int, C.c_int,
                              C.POINTER(C.c_void_p)], C.c_int),
    "inmsom_bt_ens_destroy": ([C.c_void_p], C.c_int),
    "inmsom_bt_ens_size": ([C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_longlong),
                            C.POINTER(C.c_longlong)], C.c_int),
    "inmsom_bt_ens_set_state": ([C.c_void_p, C.c_int] + [_dp] * 6, C.c_int),
    "inmsom_bt_ens_get_state": ([C.c_void_p, C.c_int] + [_dp] * 6, C.c_int),
    "inmsom_bt_ens_set_launch": ([C.c_void_p, C.c_int, C.c_int], C.c_int),
    "inmsom_bt_ens_set_substeps_per_launch": ([C.c_void_p, C.c_int], C.c_int),
    "inmsom_bt_ens_set_pipeline": ([C.c_void_p, C.c_int], C.c_int),
    "inmsom_bt_ens_get_pipeline": ([C.c_void_p, C.POINTER(C.c_int)], C.c_int),
    "inmsom_bt_ens_get_substeps_per_launch": ([C.c_void_p, C.POINTER(C.c_int)], C.c_int),
    "inmsom_bt_ens_run": ([C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int], C.c_int),
    "inmsom_bt_ens_run_timed": ([C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int, _dp], C.c_int),
    "inmsom_bt_ens_save": ([C.c_void_p], C.c_int),
    "inmsom_bt_ens_restore": ([C.c_void_p], C.c_int),
    "inmsom_bt_ens_range_check": ([C.c_void_p, C.POINTER(C.c_longlong)], C.c_int),
    "inmsom_bt_ens_volume": ([C.c_void_p, C.c_int, _dp], C.c_int),
    "inmsom_bt_cnn_create": ([C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p)], C.c_int),
    "inmsom_bt_cnn_destroy": ([C.c_void_p], C.c_int),
    "inmsom_bt_cnn_num_params": ([C.c_void_p, C.POINTER(C.c_longlong)], C.c_int),
    "inmsom_bt_cnn_set_norm": ([C.c_void_p, _fp, _fp, _fp], C.c_int),
    "inmsom_bt_cnn_set_output": ([C.c_void_p, C.c_int], C.c_int),
    "inmsom_bt_cnn_set_params": ([C.c_void_p, _fp], C.c_int),
    "inmsom_bt_cnn_get_params": ([C.c_void_p, _fp], C.c_int),
    "inmsom_bt_cnn_get_grad": ([C.c_void_p, _fp], C.c_int),
    "inmsom_bt_cnn_set_batch": ([C.c_void_p, _fp, _fp, _u8p], C.c_int),
    "inmsom_bt_cnn_forward": ([C.c_void_p, _fp], C.c_int),
    "inmsom_bt_cnn_get_activation": ([C.c_void_p, C.c_int, _fp], C.c_int),
    "inmsom_bt_cnn_loss_grad": ([C.c_void_p, _dp], C.c_int),
    "inmsom_bt_cnn_keep_backprop": ([C.c_void_p, C.c_int], C.c_int),
    "inmsom_bt_cnn_get_backprop": ([C.c_void_p, C.c_int, C.POINTER(C.c_float)], C.c_int),
    "inmsom_bt_cnn_dataset_alloc": ([C.c_void_p, C.c_int], C.c_int),
    "inmsom_bt_cnn_dataset_set": ([C.c_void_p, C.c_int, _fp, _fp, _u8p], C.c_int),
    "inmsom_bt_cnn_train_step": ([C.c_void_p, _ip, C.c_float, _dp], C.c_int),
    "inmsom_bt_cnn_eval": ([C.c_void_p, _ip, _dp], C.c_int),
    "inmsom_bt_cnn_synchronize": ([C.c_void_p], C.c_int),
    "inmsom_bt_ens_set_closure": ([C.c_void_p, C.c_void_p, C.c_double], C.c_int),
    "inmsom_bt_ens_set_closure_interval": ([C.c_void_p, C.c_int], C.c_int),
    "inmsom_bt_ens_closure_stats": ([C.c_void_p, _ip, _dp], C.c_int),
}


class BtError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where} failed with status {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """Loads libinmsom_bt.so once (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with 'make -C {PKG_DIR}/csrc' "
                "(python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def check(code: int, where: str) -> int:
    if code != OK:
        msg = lib().inmsom_bt_last_error().decode(errors="replace")
        raise BtError(code, where, msg)
    return code
