"""ctypes binding of libpauliprop.so (declarations: include/pauliprop.h).

The library is built in-tree by ``make -C paulioperators.jl_amd``.  There is
no fallback: if the library is missing, or a call fails, an exception is
raised.
"""
from __future__ import annotations

import ctypes
import os

from .truncation import PPFilter

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "_build", "libpauliprop.so")

PP_OK, PP_EARG, PP_ENOMEM, PP_EHIP, PP_EOVERFLOW, PP_EUNSUPPORTED = range(6)

EXPORTS = ("pp_create", "pp_create_typed", "pp_coeff_type", "pp_destroy", "pp_upload", "pp_download", "pp_size", "pp_capacities", "pp_evolve",
           "pp_evolve_ex", "pp_truncate", "pp_variance_ket", "pp_evolve_single", "pp_compact", "pp_lookup", "pp_lookup_noalloc", "pp_norm",
           "pp_inner_product", "pp_diff_norm", "pp_setindex", "pp_delete", "pp_sum", "pp_scale", "pp_scale_complex", "pp_scale_by_popcount", "pp_expectation_ket", "pp_check",
           "pp_copy", "pp_otimes", "pp_largest", "pp_select_diag", "pp_matrix",
           "pp_algebra",
           "pp_batch_create", "pp_batch_destroy", "pp_batch_words", "pp_batch_upload", "pp_batch_download",
           "pp_batch_sizes", "pp_batch_evolve", "pp_batch_scale_by_popcount", "pp_batch_expectation_ket",
           "pp_batch_reserve", "pp_batch_set_tiles", "pp_batch_set_timing", "pp_batch_merge_time",
           "pp_batch_set_graphs", "pp_batch_graph_stats", "pp_batch_set_resident", "pp_set_resident",
           "pp_batch_resident_stats", "pp_batch_resident_time", "pp_batch_resident_work",
           "pp_batch_set_epoch", "pp_set_epoch",
           "pp_mlp_create", "pp_mlp_destroy", "pp_mlp_param_count", "pp_mlp_set_params", "pp_mlp_get_params",
           "pp_mlp_load_data", "pp_mlp_train_epoch", "pp_mlp_gradients", "pp_mlp_predict",
           "pp_julia_powi", "pp_last_error", "pp_version")


class PPCounters(ctypes.Structure):
    """WindowCounters (spv_evolve.jl:149-161) summed over one call."""

    _fields_ = [("rotations", ctypes.c_int64), ("terms_swept", ctypes.c_int64),
                ("terms_created", ctypes.c_int64), ("merge_in", ctypes.c_int64),
                ("merge_out", ctypes.c_int64), ("early_merges", ctypes.c_int64),
                ("t_rotate", ctypes.c_double), ("t_merge", ctypes.c_double)]

    def asdict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


PP_OP_FILTER, PP_OP_ADAPTIVE, PP_OP_STOCHASTIC_COEFF, PP_OP_STOCHASTIC_SAMPLING = range(4)


class PPStrategyOp(ctypes.Structure):
    """pp_strategy_op: one step of a non-compilable boundary strategy."""

    _fields_ = [("kind", ctypes.c_int32), ("reserved", ctypes.c_int32), ("filter", PPFilter),
                ("max_terms", ctypes.c_int64), ("value", ctypes.c_double), ("seed", ctypes.c_uint64)]


PP_ALG_MUL, PP_ALG_COMMUTATOR, PP_ALG_ANTICOMMUTATOR = range(3)


class PPTerms(ctypes.Structure):
    """pp_terms: one operand of pp_algebra (n terms, W words per key)."""

    _fields_ = [("z", ctypes.POINTER(ctypes.c_uint64)), ("x", ctypes.POINTER(ctypes.c_uint64)),
                ("re", ctypes.POINTER(ctypes.c_double)), ("im", ctypes.POINTER(ctypes.c_double)),
                ("n", ctypes.c_int64), ("conj", ctypes.c_int32)]


class PPError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"pauliprop error {code}: {msg}")
        self.code = code


class ArgumentError(ValueError):
    """The reference's ArgumentError."""


class DimensionMismatch(ValueError):
    """The reference's DimensionMismatch."""


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built (run `make -C {PKG_DIR}` or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    u64p = P(ctypes.c_uint64)
    dp = P(ctypes.c_double)
    i64 = ctypes.c_int64
    L.pp_create.argtypes = [ctypes.c_int, i64, i64, P(vp)]
    L.pp_create_typed.argtypes = [ctypes.c_int, i64, i64, ctypes.c_int, P(vp)]
    L.pp_coeff_type.argtypes = [vp, P(ctypes.c_int)]
    L.pp_destroy.argtypes = [vp]
    L.pp_upload.argtypes = [vp, u64p, u64p, dp, i64, ctypes.c_int]
    L.pp_download.argtypes = [vp, u64p, u64p, dp, i64, P(i64)]
    L.pp_size.argtypes = [vp, P(i64)]
    L.pp_capacities.argtypes = [vp, P(i64), P(i64)]
    L.pp_evolve.argtypes = [vp, u64p, u64p, dp, dp, i64, i64, P(PPFilter), P(PPFilter), P(PPCounters)]
    L.pp_evolve_ex.argtypes = [vp, u64p, u64p, dp, dp, i64, i64, P(PPFilter), P(PPStrategyOp), ctypes.c_int,
                               P(PPFilter), u64p, dp, dp, P(PPCounters)]
    L.pp_truncate.argtypes = [vp, P(PPStrategyOp), ctypes.c_int, u64p, dp, dp]
    L.pp_variance_ket.argtypes = [vp, u64p, dp]
    L.pp_evolve_single.argtypes = [vp, u64p, u64p, ctypes.c_double, ctypes.c_double]
    L.pp_lookup.argtypes = [vp, u64p, u64p, i64, dp, P(ctypes.c_int32)]
    L.pp_lookup_noalloc.argtypes = [vp, u64p, u64p, i64, dp, P(ctypes.c_int32)]
    L.pp_norm.argtypes = [vp, ctypes.c_double, dp]
    L.pp_setindex.argtypes = [vp, u64p, u64p, dp, i64]
    L.pp_sum.argtypes = [vp, vp]
    L.pp_scale.argtypes = [vp, ctypes.c_double]
    L.pp_scale_complex.argtypes = [vp, ctypes.c_double, ctypes.c_double]
    L.pp_delete.argtypes = [vp, u64p, u64p, i64]
    L.pp_inner_product.argtypes = [vp, vp, dp]
    L.pp_diff_norm.argtypes = [vp, vp, dp]
    L.pp_compact.argtypes = [vp, P(PPFilter)]
    L.pp_scale_by_popcount.argtypes = [vp, ctypes.c_int, u64p, dp, dp, dp]
    L.pp_expectation_ket.argtypes = [vp, u64p, dp]
    L.pp_check.argtypes = [vp, P(ctypes.c_int)]
    L.pp_copy.argtypes = [vp, vp]
    L.pp_otimes.argtypes = [vp, vp, vp]
    L.pp_largest.argtypes = [vp, ctypes.c_int, u64p, u64p, dp, P(i64)]
    L.pp_select_diag.argtypes = [vp, ctypes.c_int]
    L.pp_matrix.argtypes = [vp, dp]
    L.pp_algebra.argtypes = [vp, ctypes.c_int, ctypes.c_int, P(PPTerms), P(PPTerms), u64p, u64p, dp, dp, i64, P(i64)]
    ip = P(ctypes.c_int)
    L.pp_batch_create.argtypes = [ctypes.c_int, ip, P(i64), P(i64), P(vp)]
    L.pp_batch_destroy.argtypes = [vp]
    L.pp_batch_words.argtypes = [vp, ip]
    L.pp_batch_upload.argtypes = [vp, ctypes.c_int, u64p, u64p, dp, i64, ctypes.c_int]
    L.pp_batch_download.argtypes = [vp, ctypes.c_int, u64p, u64p, dp, i64, P(i64)]
    L.pp_batch_sizes.argtypes = [vp, P(i64)]
    L.pp_batch_evolve.argtypes = [vp, i64, u64p, u64p, dp, dp, P(ctypes.c_uint8), P(PPFilter), P(PPFilter),
                                  P(PPCounters)]
    L.pp_batch_scale_by_popcount.argtypes = [vp, ip, u64p, dp]
    L.pp_batch_expectation_ket.argtypes = [vp, u64p, dp]
    L.pp_batch_reserve.argtypes = [vp, P(i64)]
    L.pp_batch_set_tiles.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.pp_batch_set_timing.argtypes = [vp, ctypes.c_int]
    L.pp_batch_set_graphs.argtypes = [vp, ctypes.c_int]
    L.pp_batch_graph_stats.argtypes = [vp, P(i64), P(i64)]
    L.pp_batch_merge_time.argtypes = [vp, dp, P(i64), ctypes.c_int]
    L.pp_batch_set_resident.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.pp_set_resident.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.pp_batch_resident_stats.argtypes = [vp, P(i64), P(i64), P(i64)]
    L.pp_batch_resident_time.argtypes = [vp, dp, ctypes.c_int]
    L.pp_batch_resident_work.argtypes = [vp, P(i64), P(i64), P(i64), P(i64), ctypes.c_int]
    L.pp_batch_set_epoch.argtypes = [vp, ctypes.c_uint32]
    L.pp_set_epoch.argtypes = [vp, ctypes.c_uint32]
    fp = P(ctypes.c_float)
    L.pp_mlp_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, P(vp)]
    L.pp_mlp_destroy.argtypes = [vp]
    L.pp_mlp_param_count.argtypes = [vp, P(i64)]
    L.pp_mlp_set_params.argtypes = [vp, fp]
    L.pp_mlp_get_params.argtypes = [vp, fp]
    L.pp_mlp_load_data.argtypes = [vp, fp, fp, i64]
    L.pp_mlp_train_epoch.argtypes = [vp, P(ctypes.c_int32), ctypes.c_int, ctypes.c_float, fp]
    L.pp_mlp_gradients.argtypes = [vp, fp, fp, ctypes.c_int, fp, fp]
    L.pp_mlp_predict.argtypes = [vp, fp, i64, fp, dp, dp, dp]
    L.pp_julia_powi.argtypes = [ctypes.c_double, i64]
    L.pp_julia_powi.restype = ctypes.c_double
    L.pp_last_error.restype = ctypes.c_char_p
    L.pp_version.restype = ctypes.c_char_p
    for name in EXPORTS:
        if name not in ("pp_julia_powi", "pp_last_error", "pp_version"):
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def check(rc: int):
    if rc != PP_OK:
        msg = lib().pp_last_error().decode(errors="replace")
        if rc == PP_EARG:
            raise ArgumentError(msg)
        raise PPError(rc, msg)


def u64p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
