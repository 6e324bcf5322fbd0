"""The C ABI library loads and exports every entry point include/vsim.h
declares (no device calls: this runs without a GPU)."""
import ctypes
import os
import re

from vsim import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "vsim.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*\**(vs_\w+)\(", text, re.M)))


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(native.LIB_PATH)
    names = declared()
    assert len(names) >= 29
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing
    assert set(names) == set(native.EXPORTS)


def test_kernel_ids_follow_the_reference_pass_labels():
    lib = native.load()
    labels = ["sim-predict", "sim-finalize", "sim-elastic-cells", "sim-contact-bin", "sim-contact-sort",
              "sim-contact-collide", "sim-contact-settle", "sim-drive-hard", "sim-drive-compliant"]
    assert [lib.vs_kernel_id(l.encode()) for l in labels] == list(range(9))
    assert lib.vs_kernel_id(b"sim-unknown") < 0


def test_struct_sizes_match_the_header():
    assert ctypes.sizeof(native.VsBody) == 80
    assert ctypes.sizeof(native.VsSceneParams) == 160
    assert ctypes.sizeof(native.VsLaunch) == 32 + 80
