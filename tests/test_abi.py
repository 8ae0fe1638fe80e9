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
    assert L.mts_batch_run(None, 1, None, None, None, None) == -1
    assert L.mts_batch_run(None, 0, None, None, None, None) == 0  # nothing to run
    assert L.mts_server_create(None, 1 << 30, 4, 2, None) == -1
    assert L.mts_server_submit(None, None, None) == -1
    assert L.mts_server_wait(None, 0) == -1
    L.mts_server_destroy(None)


def test_request_and_result_layout_match_the_header(mts, tmp_path):
    """ctypes mirrors of mts_request / mts_result agree with the C compiler's layout."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        import pytest
        pytest.skip("no C compiler")
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "mts.h"\n'
        'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(mts_request), '
        'offsetof(mts_request, tag), offsetof(mts_request, precision), sizeof(mts_result), '
        'offsetof(mts_result, message), offsetof(mts_result, argmax), '
        'offsetof(mts_result, footprint));return 0;}\n')
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True,
                                          text=True).stdout.split()]
    want = [ctypes.sizeof(mts.Request), mts.Request.tag.offset, mts.Request.precision.offset,
            ctypes.sizeof(mts.Result),
            mts.Result.message.offset, mts.Result.argmax.offset, mts.Result.footprint.offset]
    assert got == want


def test_oracle_library_exports(oracle):
    lib = oracle.lib()
    for n in ("mts_oracle_init", "mts_oracle_run", "mts_oracle_norms", "mts_oracle_tdma"):
        assert hasattr(lib, n)
