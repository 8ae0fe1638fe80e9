"""The C-ABI library loads and exports every entry point include/mts.h declares."""
import ctypes
import os
import re

from conftest import PKG_DIR, ROOT

HEADER = os.path.join(ROOT, "include", "mts.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mts_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("mts_setup_create", "mts_instance_create", "mts_step", "mts_run",
              "mts_error_norms", "mts_tdma_batched", "mts_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol(mts):
    lib = ctypes.CDLL(os.path.join(PKG_DIR, "_build", "libmts.so"))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_version_and_error_string(mts):
    assert "gfx950" in mts.version()
    assert isinstance(mts.lib().mts_last_error(), bytes)


def test_null_arguments_are_errors_not_crashes(mts):
    L = mts.lib()
    assert L.mts_setup_create(None, None, None, 1, 2, None) == -1
    assert b"null" in L.mts_last_error()
    assert L.mts_step(None, 0.1) == -1
    assert L.mts_ctx_create(0, None) == -1


def test_oracle_library_exports(oracle):
    lib = oracle.lib()
    for n in ("mts_oracle_init", "mts_oracle_run", "mts_oracle_norms", "mts_oracle_tdma"):
        assert hasattr(lib, n)
