"""The C-ABI library loads and exports every entry point include/ot/*.h declares
(CPU only: no compute call needs a GPU here)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER_DIR = os.path.join(ROOT, "include", "ot")


def declared_functions():
    names = set()
    for fn in os.listdir(HEADER_DIR):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(HEADER_DIR, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(ot_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_functions()
    for required in ("ot_derive", "ot_plan_create", "ot_run", "ot_launch", "ot_plan_destroy",
                     "ot_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol(otlib):
    exported = subprocess.run(["nm", "-D", "--defined-only", otlib.LIB_PATH], capture_output=True,
                              text=True, check=True).stdout
    symbols = {line.split()[-1] for line in exported.splitlines() if line.strip()}
    missing = sorted(declared_functions() - symbols)
    assert not missing, missing


def test_binding_covers_every_declared_symbol(otlib):
    bound = {name for name, _, _ in otlib.SIGNATURES}
    missing = sorted(declared_functions() - bound)
    assert not missing, missing


def test_version_and_catalogue(otlib):
    assert b"gfx950" in otlib.lib().ot_version()
    for i, name in enumerate(otlib.STENCILS):
        info = otlib.stencil_info(i)
        assert info.name.decode() == name
        import ctypes as C
        sid = C.c_int32(-1)
        otlib.check(otlib.lib().ot_stencil_from_name(name.encode(), C.byref(sid)))
        assert sid.value == i
    with pytest.raises(otlib.OtError):
        otlib.stencil_info(99)


def test_error_message_reported(otlib):
    with pytest.raises(otlib.OtError) as ex:
        otlib.derive(otlib.make_params("jacobi-2d", T=0))
    assert "time tile" in str(ex.value)
