"""CPU checks of the drop-in boundary: the C-ABI libraries load and export
every entry point include/hsps_dev.h and include/hsps_serve.h declare;
host-side argument errors come back as status codes (no GPU calls made)."""
import os
import re

import pytest

from conftest import ROOT


def _declared(header="hsps_dev.h"):
    with open(os.path.join(ROOT, "include", header)) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hsps_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = _declared()
    for must in ["hsps_ctx_create", "hsps_csr_create", "hsps_spmv_f64", "hsps_cg_solve", "hsps_fgmres_solve",
                 "hsps_precond_create", "hsps_batch_solve", "hsps_features"]:
        assert must in names


def test_library_exports_every_declared_symbol(H):
    L = H.lib()
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing
    # and the Python binding covers exactly the declared surface
    assert sorted(H.SIGNATURES) == _declared()


def test_null_arguments_are_status_codes(H):
    L = H.lib()
    assert L.hsps_ctx_create(0, None) == -1
    assert b"null" in L.hsps_last_error().lower()
    p = H.default_params()
    assert p.tol == 1e-8 and p.max_iter == 2000 and p.restart_size == 50
    assert abs(p.amg_omega - 0.67) < 1e-15 and p.amg_strong == 0.25 and p.drop_tol == 1e-4
    st = H.Stats()
    assert L.hsps_cg_solve(None, None, None, None, None, st) == -1


def test_serve_library_exports_every_declared_symbol(H):
    L = H.serve_lib()
    names = _declared("hsps_serve.h")
    assert {"hsps_server_create", "hsps_server_submit", "hsps_server_wait", "hsps_server_poll"} <= set(names)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(H.SERVE_SIGNATURES) == names


def test_serve_null_arguments_are_status_codes(H):
    L = H.serve_lib()
    assert L.hsps_server_create(0, 1, 1e9, 8, 1000, 0, None) == -1
    assert b"null" in L.hsps_server_last_error().lower()
    t = H.C.c_longlong()
    assert L.hsps_server_submit(None, 4, None, None, None, None, None, H.C.byref(t)) == -1
