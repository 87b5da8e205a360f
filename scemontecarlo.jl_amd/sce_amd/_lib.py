"""ctypes binding of ``libsce_hip.so`` (include/sce_hip.h).

The product path has no fallback: if the library is missing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libsce_hip.so")

SCE_OK = 0
ERRORS = {-1: "SCE_EINVAL", -2: "SCE_EDIM", -3: "SCE_ERANGE_COUNTER", -4: "SCE_EHIP",
          -5: "SCE_ENOMEM", -6: "SCE_ELAYOUT", -7: "SCE_ENODEV"}


class SceError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class ArgumentError(SceError, ValueError):
    """SCE_EINVAL — the reference's ArgumentError."""


class DimensionMismatch(SceError, ValueError):
    """SCE_EDIM — the reference's DimensionMismatch."""


i32p = C.POINTER(C.c_int32)
i8p = C.POINTER(C.c_int8)
u8p = C.POINTER(C.c_uint8)
f64p = C.POINTER(C.c_double)


class TablesDesc(C.Structure):
    _fields_ = [
        ("n_cell_atoms", C.c_int32), ("dims", C.c_int32 * 3),
        ("n_sites", C.c_int32), ("lmax", C.c_int32), ("nlm", C.c_int32),
        ("n_terms", C.c_int32), ("n_inst", C.c_int32), ("n_active", C.c_int32),
        ("n_colors", C.c_int32), ("n_prog", C.c_int32),
        ("n_inst_sites", C.c_int64), ("n_sent", C.c_int64), ("n_sfac", C.c_int64),
        ("n_eent", C.c_int64), ("n_efac", C.c_int64),
        ("inst_term", i32p), ("inst_ptr", i32p), ("inst_sites", i32p),
        ("site_ptr", i32p), ("site_inst", i32p), ("site_slot", i8p),
        ("site_has_l1", u8p), ("site_active", u8p),
        ("site_prog", i32p), ("sprog_ptr", i32p), ("sent_w", f64p),
        ("sent_tgt", i32p), ("sfac_ptr", i32p), ("sfac_row", i32p), ("sfac_slot", i8p),
        ("site_col", i32p), ("site_col2", i32p), ("pent_row", i32p), ("pent_row2", i32p),
        ("eprog_ptr", i32p), ("eent_w", f64p), ("efac_ptr", i32p), ("efac_row", i32p),
        ("term_coef", f64p), ("color_ptr", i32p), ("color_sites", i32p),
    ]


class ChainInfo(C.Structure):
    _fields_ = [("energy", C.c_double), ("step", C.c_double), ("max_drift", C.c_double),
                ("last_drift", C.c_double), ("sweep_index", C.c_int64),
                ("acc_metro", C.c_int64), ("att_metro", C.c_int64), ("acc_or", C.c_int64),
                ("att_or", C.c_int64), ("seed", C.c_uint64), ("acc_stage1", C.c_int64)]


# (name, length-expression, dtype) of every array field of TablesDesc
ARRAY_FIELDS = [
    ("inst_term", "n_inst", np.int32), ("inst_ptr", "n_inst+1", np.int32),
    ("inst_sites", "n_inst_sites", np.int32), ("site_ptr", "n_sites+1", np.int32),
    ("site_inst", "n_inst_sites", np.int32), ("site_slot", "n_inst_sites", np.int8),
    ("site_has_l1", "n_sites", np.uint8), ("site_active", "n_sites", np.uint8),
    ("site_prog", "n_inst_sites", np.int32), ("sprog_ptr", "n_prog+1", np.int32),
    ("sent_w", "n_sent", np.float64), ("sent_tgt", "n_sent", np.int32),
    ("sfac_ptr", "n_sent+1", np.int32), ("sfac_row", "n_sfac", np.int32),
    ("sfac_slot", "n_sfac", np.int8), ("site_col", "n_inst_sites", np.int32),
    ("site_col2", "n_inst_sites", np.int32), ("pent_row", "n_sent", np.int32),
    ("pent_row2", "n_sent", np.int32), ("eprog_ptr", "n_terms+1", np.int32),
    ("eent_w", "n_eent", np.float64), ("efac_ptr", "n_eent+1", np.int32),
    ("efac_row", "n_efac", np.int32), ("term_coef", "n_terms", np.float64),
    ("color_ptr", "n_colors+1", np.int32), ("color_sites", "n_active", np.int32),
]

_lib = None

# every symbol include/sce_hip.h declares
EXPORTS = [
    "sce_version", "sce_last_error", "sce_device_count", "sce_tiled_build", "sce_tiled_desc",
    "sce_tiled_free", "sce_host_zrows", "sce_host_total_energy", "sce_tables_create",
    "sce_tables_destroy", "sce_tables_layout_info", "sce_layout_check", "sce_tables_position_tables",
    "sce_position_tables_host", "sce_chain_create", "sce_chain_destroy",
    "sce_chain_info_get", "sce_chain_set_step", "sce_chain_reset_counters", "sce_download",
    "sce_upload", "sce_swap_payload", "sce_metropolis_sweep", "sce_overrelaxation_sweep",
    "sce_run_sweeps", "sce_renormalize", "sce_chain_total_energy", "sce_chain_set_timing", "sce_chain_set_persistent",
    "sce_chain_timing", "sce_test_philox", "sce_test_detmath", "sce_test_zlm",
    "sce_test_proposal", "sce_test_walk_delta", "sce_test_walk_coeffs", "sce_test_reduce",
    "sce_batch_create", "sce_batch_destroy", "sce_batch_run", "sce_batch_download",
    "sce_batch_info", "sce_batch_set_step", "sce_batch_reset_counters", "sce_batch_swap_stats",
    "sce_batch_set_graph", "sce_batch_set_persistent", "sce_batch_set_shared_walk", "sce_batch_set_site_pairs", "sce_batch_renormalize", "sce_batch_set_counters",
    "sce_batch_get_exchange", "sce_batch_set_exchange",
    "sce_sur_create", "sce_sur_destroy", "sce_sur_info", "sce_sur_get_params",
    "sce_sur_set_params", "sce_sur_set_int8", "sce_sur_get_qparams", "sce_sur_harvest", "sce_sur_train", "sce_sur_grad", "sce_sur_eval",
    "sce_sur_features", "sce_da_run_sweeps", "sce_chain_set_counters", "sce_chain_time_walk",
    "sce_chain_set_graph", "sce_chain_set_drift", "sce_chain_binner_reset", "sce_chain_binner_off", "sce_chain_binner_get",
    "sce_energy_gradient", "sce_energy_gradient_noalloc", "sce_chain_energy_gradient", "sce_chain_time_gradient",
    "sce_test_mfma", "sce_test_chain_plan",
]


def lib():
    """Load libsce_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() "
                              "(hipcc, gfx950) first — there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        P, i32, i64, u64, d = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64, C.c_double
        sig = {
            "sce_version": [], "sce_last_error": [C.c_char_p, C.c_size_t],
            "sce_device_count": [P],
            "sce_tiled_build": [i32, i32, P, P, P, P, P, P, P, P, P, P],
            "sce_tiled_desc": [P, P], "sce_tiled_free": [P],
            "sce_host_zrows": [i32, i64, P, P], "sce_host_total_energy": [P, P, P],
            "sce_tables_create": [P, P], "sce_tables_destroy": [P],
            "sce_tables_layout_info": [P, P], "sce_layout_check": [P, P],
            "sce_tables_position_tables": [P, P, P, P, P], "sce_position_tables_host": [P, P, P, P, P],
            "sce_chain_create": [P, P, P, d, u64, d, P], "sce_chain_destroy": [P],
            "sce_chain_info_get": [P, P], "sce_chain_set_step": [P, d],
            "sce_chain_reset_counters": [P], "sce_download": [P, P, P, P],
            "sce_upload": [P, P, P, d], "sce_swap_payload": [P, P],
            "sce_metropolis_sweep": [P, d, i32, P],
            "sce_overrelaxation_sweep": [P, d, i32, i32, P],
            "sce_run_sweeps": [P, d, i64, i32, i64, i64, i32, P, i64, P],
            "sce_renormalize": [P, P], "sce_chain_total_energy": [P, P],
            "sce_chain_set_timing": [P, i32], "sce_chain_set_persistent": [P, i32], "sce_chain_timing": [P, P, P, P],
            "sce_test_philox": [i64, P, P, P], "sce_test_detmath": [i64, i32, P, P],
            "sce_test_zlm": [i64, i32, P, P],
            "sce_test_proposal": [i64, u64, P, P, P, d, P, P],
            "sce_test_walk_delta": [P, P, i64, P, P, i32, P],
            "sce_test_walk_coeffs": [P, P, i64, P, i32, P],
            "sce_test_reduce": [i64, P, P],
            "sce_batch_create": [i32, P, P, P, P, P, P, P, u64, P],
            "sce_batch_destroy": [P],
            "sce_batch_run": [P, i64, i32, i64, i64, i64, i64, i64, P, i64, P],
            "sce_batch_download": [P, i32, P, P, P], "sce_batch_info": [P, i32, P],
            "sce_batch_set_step": [P, i32, d], "sce_batch_reset_counters": [P],
            "sce_batch_swap_stats": [P, P, P], "sce_batch_set_graph": [P, i32], "sce_batch_set_persistent": [P, i32], "sce_batch_set_shared_walk": [P, i32], "sce_batch_set_site_pairs": [P, i32],
            "sce_batch_renormalize": [P],
            "sce_batch_set_counters": [P, i32, i64, i64, i64, i64, i64, d],
            "sce_batch_get_exchange": [P, P, P], "sce_batch_set_exchange": [P, i64, i32, P, P],
            "sce_sur_create": [P, u64, P], "sce_sur_destroy": [P],
            "sce_sur_info": [P, i32, P, P, P, P], "sce_sur_get_params": [P, i32, P],
            "sce_sur_set_params": [P, i32, P, d], "sce_sur_set_int8": [P, i32], "sce_sur_get_qparams": [P, i32, P], "sce_sur_harvest": [P, P, d, i64],
            "sce_sur_train": [P, i32, i32, i32, d, u64, P],
            "sce_sur_grad": [P, i32, i32, P, P, P, P, P],
            "sce_sur_eval": [P, i32, i32, P, P, P],
            "sce_sur_features": [P, P, i32, P, P, P, P],
            "sce_da_run_sweeps": [P, P, d, i64, i64, i64, P, i64, P],
            "sce_chain_set_counters": [P, i64, i64, i64, i64, i64],
            "sce_chain_time_walk": [P, i32, i32, P, P], "sce_chain_set_graph": [P, i32],
            "sce_chain_set_drift": [P, d],
            "sce_chain_binner_reset": [P, i32, i32], "sce_chain_binner_off": [P],
            "sce_chain_binner_get": [P, P, P, P, P, P, P, i32],
            "sce_energy_gradient": [P, P, P, i32], "sce_energy_gradient_noalloc": [P, P, P, i32],
            "sce_chain_energy_gradient": [P, P, i32],
            "sce_chain_time_gradient": [P, i32, i32, P],
            "sce_test_mfma": [i64, P, P, P, P],
            "sce_test_chain_plan": [i32, P, i32, i32, i32, i32, P, i32, P, P],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.restype = C.c_int32
            f.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    lib().sce_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(rc: int) -> None:
    if rc == SCE_OK:
        return
    msg = last_error()
    if rc == -1:
        raise ArgumentError(rc, msg)
    if rc == -2:
        raise DimensionMismatch(rc, msg)
    raise SceError(rc, msg)


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def device_count() -> int:
    n = C.c_int32(0)
    lib().sce_device_count(C.byref(n))
    return n.value
