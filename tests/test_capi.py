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


MIRRORS = {"FrameArgs": "gray_frame_args", "Knobs": "gray_knobs", "Params": "gray_params", "View": "gray_view",
           "RenderArgs": "gray_render_args", "Grads": "gray_grads", "Adam": "gray_adam",
           "LossArgs": "gray_loss_args"}


@pytest.mark.skipif(not os.path.exists(LIB), reason="libgray_hip.so not built")
def test_ctypes_structures_match_the_c_layout(tmp_path):
    """Size and every field offset of the binding's structures, against the
    header compiled by the host C compiler."""
    import shutil
    import subprocess

    from gray_amd import _lib

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "gray_hip.h"', "int main(void) {"]
    for py, c in MIRRORS.items():
        cls = getattr(_lib, py)
        lines.append(f'printf("{py} size %zu\\n", sizeof({c}));')
        for name, *_ in cls._fields_:
            lines.append(f'printf("{py} {name} %zu\\n", offsetof({c}, {name}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {(p[0], p[1]): int(p[2]) for p in (l.split() for l in out if l)}
    for py in MIRRORS:
        cls = getattr(_lib, py)
        assert got[(py, "size")] == ctypes.sizeof(cls), py
        for name, *_ in cls._fields_:
            assert got[(py, name)] == getattr(cls, name).offset, (py, name)
