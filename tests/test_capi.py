"""The C ABI library loads on a machine without a GPU and exports every function that
include/spand.h declares; argument and state errors come back as status codes."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import spand_public_amd as S
from spand_public_amd import _lib
from spand_public_amd.problems import linspace_nd, neglapl, symmetric_graph

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "spand.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(spand_[A-Za-z_0-9]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    L = _lib.lib()
    syms = declared_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(L, s), s
    # and the Python binding declares a signature for each of them
    assert set(syms) == set(_lib.SIGNATURES)


def test_status_codes_without_gpu():
    L = _lib.lib()
    h = C.c_void_p()
    assert L.spand_create(0, C.byref(h)) == 1  # SPAND_ERR_ARG
    assert L.spand_create(3, C.byref(h)) == 0
    assert L.spand_factorize(h) == 5  # SPAND_ERR_STATE: factorize before assemble
    assert L.spand_set_symm_kind(h, 1) == 8  # SPAND_ERR_UNSUPPORTED (SYM / LDLT not built)
    assert L.spand_set_tol(h, -1.0) == 1
    assert b"tol" in L.spand_last_error(h)
    A = neglapl(5, 2)
    # partition without coordinates: only geometric partitioning is available
    assert L.spand_partition(h, A.n, A.colptr, A.rowval, None) == 8
    L.spand_destroy(h)


def test_python_mirror_errors():
    t = S.Tree(3)
    with pytest.raises(S.SpandError):
        t.factorize()
    with pytest.raises(S.SpandError):
        S.Tree(0)
    A, X = neglapl(5, 2), linspace_nd(5, 2)
    t.set_use_geo(True)
    t.set_Xcoo(X)
    t.partition(symmetric_graph(A))
    t.assemble(A)
    assert t.get_N() == 25
    with pytest.raises(S.SpandError):
        t.solve(np.ones(25))  # solve before factorize


def test_symm_kind_is_fixed_by_assembly():
    """assemble() builds lower blocks only for SPD, so the kind cannot change after it."""
    A, X = neglapl(5, 2), linspace_nd(5, 2)
    t = S.Tree(3)
    t.set_use_geo(True)
    t.set_Xcoo(X)
    t.set_symm_kind("GEN")
    t.partition(symmetric_graph(A))
    t.assemble(A)
    t.set_symm_kind("GEN")  # unchanged: accepted
    with pytest.raises(S.SpandError) as e:
        t.set_symm_kind("SPD")
    assert e.value.status == 5  # SPAND_ERR_STATE
    # a new assembly of the same tree keeps working with its kind
    t.assemble(A)


def test_device_values_need_an_assembly_and_the_pattern_size():
    L = _lib.lib()
    A, X = neglapl(5, 2), linspace_nd(5, 2)
    t = S.Tree(3)
    # before assemble: state error, nothing read
    assert L.spand_assemble_values_device(t._h, None, A.nnz, None) == 5
    t.set_use_geo(True)
    t.set_Xcoo(X)
    t.partition(symmetric_graph(A))
    t.assemble(A)
    # wrong entry count: argument error, nothing read
    assert L.spand_assemble_values_device(t._h, None, A.nnz + 1, None) == 1
    assert b"nnz" in L.spand_last_error(t._h)


def test_shape_log_and_status_flags_without_gpu():
    L = _lib.lib()
    h = C.c_void_p()
    assert L.spand_create(3, C.byref(h)) == 0
    n = C.c_longlong(-1)
    assert L.spand_get_flops_log(h, C.byref(n), None) == 0 and n.value == 0  # nothing recorded yet
    assert L.spand_get_flops_log(h, None, None) == 1  # SPAND_ERR_ARG
    assert b"count" in L.spand_last_error(h)
    assert L.spand_set_monitor_flops(h, 1) == 0
    f = C.c_uint(7)
    assert L.spand_get_status_flags(h, C.byref(f)) == 0 and f.value == 0
    assert L.spand_get_status_flags(h, None) == 1
    L.spand_destroy(h)
    t = S.Tree(3)
    assert t.flops_log().shape == (0, 5)
    assert t.status_flags() == {"step_throttle_gave_up": False}
