"""The C ABI: the library loads, exports every entry point that
include/bonxai_gpu.h declares, the Python layer binds all of them, and a
machine without a GPU gets an error (no CPU fallback)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bonxai_gpu.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w\s\*]*?\b(bxg_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "bxg_insert_point_cloud" in names and "bxg_lookup" in names
    assert len(names) > 50


def test_library_exports_every_declared_symbol():
    import bonxai_amd as bx
    out = subprocess.run(["nm", "-D", "--defined-only", bx.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\sT\s(bxg_\w+)", out))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing


def test_python_layer_binds_every_symbol():
    import bonxai_amd as bx
    bound = {name for name, _, _ in bx.ABI}
    missing = [n for n in declared() if n not in bound]
    assert not missing, missing
    lib = bx.lib()
    for n in declared():
        assert hasattr(lib, n)


def test_status_strings():
    import bonxai_amd as bx
    assert bx.lib().bxg_status_string(bx.ERR_OUT_OF_POOL).decode().startswith("node pool")


def _has_gpu():
    return os.path.exists("/dev/kfd")


@pytest.mark.skipif(_has_gpu(), reason="a GPU is present")
def test_no_gpu_fails_loudly():
    import bonxai_amd as bx
    with pytest.raises(bx.BonxaiError):
        bx.Context(0)
