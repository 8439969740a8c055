"""The C-ABI library loads and exports every entry point include/afq.h declares (no GPU calls)."""
import ctypes
import os
import re

import pytest

from safire_amd import afqmc as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "afq.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(afq_\w+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = header_functions()
    for required in ("afq_ham_create", "afq_vbias", "afq_vhs", "afq_energy",
                     "afq_det_mixed_density_matrix", "afq_det_propagate",
                     "afq_det_orthogonalize", "afq_step", "afq_orthogonalize"):
        assert required in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(A.LIB_PATH):
        pytest.fail("libsafire_afq.so not built: run __graft_entry__.build()")
    L = ctypes.CDLL(A.LIB_PATH)
    missing = [n for n in header_functions() if not hasattr(L, n)]
    assert not missing, missing
    assert set(header_functions()) == set(A.SIGNATURES), "Python signatures out of date"


def test_host_only_entry_points_work_without_gpu():
    p = A.default_params()
    assert p.taylor_n == 6 and p.vbias_bound == 50.0 and p.hybrid == 1
    assert A.lib().afq_version() >= 1
    # a null handle is rejected with an error code, not a crash
    assert A.lib().afq_ham_reserve(None, 4) == 4
    assert b"null" in A.lib().afq_last_error()
