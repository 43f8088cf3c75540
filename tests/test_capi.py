"""The C-ABI library loads and exports every entry point include/gray_hip.h
declares (CPU: no compute calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gray_hip.h")
LIB = os.path.join(ROOT, "gray_amd", "libgray_hip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(gray_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for required in ("gray_ctx_create", "gray_bvh_update", "gray_render_forward", "gray_render_backward",
                     "gray_sh_forward", "gray_sh_backward", "gray_adam_step", "gray_debug_hitlists"):
        assert required in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libgray_hip.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    lib.gray_abi_version.restype = ctypes.c_int
    m = re.search(r"#define GRAY_ABI_VERSION (\d+)", open(HEADER).read())
    assert lib.gray_abi_version() == int(m.group(1))
    lib.gray_last_error.restype = ctypes.c_char_p
    assert isinstance(lib.gray_last_error(), bytes)


@pytest.mark.skipif(not os.path.exists(LIB), reason="libgray_hip.so not built")
def test_python_binding_matches_header():
    from gray_amd import _lib

    assert sorted(_lib.EXPORTED) == declared_functions()
    # argument validation happens before any device work
    rc = _lib.LIB.gray_bvh_update(None, None, None, 1, None)
    assert rc == _lib.GRAY_ERR_INVALID
    assert b"null" in _lib.LIB.gray_last_error()
