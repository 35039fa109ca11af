"""CPU checks of the C ABI: the library loads, exports every function that
include/pauliprop.h declares, and its host-side helpers agree with the
oracle.  No device calls."""
import ctypes
import os
import re

import pytest

import pyoracle as po
from pauliprop import _lib
from pauliprop.truncation import PPFilter

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    hdr = open(os.path.join(ROOT, "include", "pauliprop.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(pp_[a-z_0-9]+)\s*\(", hdr)))


def test_header_declares_the_python_exports():
    assert set(declared_functions()) == set(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    for name in declared_functions():
        assert hasattr(L, name), name


def test_filter_struct_is_mergefilter_layout():
    # spv_kernels.jl:33-42: 3 x Int64 then 5 x Float64, 64 bytes, no padding
    assert ctypes.sizeof(PPFilter) == 64
    assert PPFilter.thresh.offset == 24
    assert PPFilter.dthresh_xw.offset == 56


def test_julia_powi_matches_oracle():
    L = _lib.lib()
    for lam in (1 - 4 * 1e-3 / 3, 1 - 2 * 5e-4, 0.37, 0.999, 1.0, 0.0, -0.5):
        for k in range(0, 129):
            assert L.pp_julia_powi(lam, k) == po.julia_powi(lam, k)


def test_version_string():
    assert _lib.lib().pp_version().startswith(b"pauliprop")


def test_strategy_op_struct_layout():
    # pp_strategy_op: int32 kind, int32 reserved, pp_filter (64 B), int64, double, uint64
    assert ctypes.sizeof(_lib.PPStrategyOp) == 96
    assert _lib.PPStrategyOp.filter.offset == 8
    assert _lib.PPStrategyOp.max_terms.offset == 72
    assert _lib.PPStrategyOp.seed.offset == 88
    hdr = open(os.path.join(ROOT, "include", "pauliprop.h")).read()
    assert re.search(r"PP_OP_FILTER = 0, PP_OP_ADAPTIVE = 1, PP_OP_STOCHASTIC_COEFF = 2, "
                     r"PP_OP_STOCHASTIC_SAMPLING = 3", hdr)
    assert (_lib.PP_OP_FILTER, _lib.PP_OP_ADAPTIVE, _lib.PP_OP_STOCHASTIC_COEFF,
            _lib.PP_OP_STOCHASTIC_SAMPLING) == (0, 1, 2, 3)
