"""The C ABI: the map library loads, exports every entry point that
include/bonxai_gpu.h declares (and the probe library every one of
include/bonxai_gpu_probes.h, none of which the map library carries), the
Python layer binds all of them, and a machine without a GPU gets an error
(no CPU fallback)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bonxai_gpu.h")
PROBE_HEADER = os.path.join(ROOT, "include", "bonxai_gpu_probes.h")


def declared(header=HEADER):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w\s\*]*?\b(bxg_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "bxg_insert_point_cloud" in names and "bxg_lookup" in names
    assert len(names) > 50


def exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                         check=True).stdout
    return set(re.findall(r"\sT\s(bxg_\w+)", out))


def test_library_exports_every_declared_symbol():
    import bonxai_amd as bx
    missing = [n for n in declared() if n not in exported(bx.LIB_PATH)]
    assert not missing, missing


def test_probes_live_in_their_own_library():
    import bonxai_amd as bx
    probes = declared(PROBE_HEADER)
    assert probes == ["bxg_copy_probe", "bxg_fp64_probe", "bxg_gather_probe"]
    assert not set(probes) & set(declared())
    assert not set(probes) & exported(bx.LIB_PATH)
    missing = [n for n in probes if n not in exported(bx.PROBE_LIB_PATH)]
    assert not missing, missing
    assert {name for name, _, _ in bx.PROBE_ABI} == set(probes)
    handle = bx.probes()
    for n in probes:
        assert hasattr(handle, n)


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
