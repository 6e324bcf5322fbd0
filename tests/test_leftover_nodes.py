"""The node tiles of the first and last colour-pair launches index a body's
leftover nodes through vsim::Leftover (kernels.h): entry i of the list maps to
a node id by a closed form. This builds leftover_check.cpp, which compares
that map with brute force over every lattice of 1..21 cells a side and each
row class (exact list, node-id order, every id inside the body), on the CPU."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(shutil.which(HIPCC) is None, reason="hipcc not found")
def test_leftover_enumeration_matches_brute_force(tmp_path):
    exe = str(tmp_path / "leftover_check")
    subprocess.run([HIPCC, "-O1", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    os.path.join(HERE, "leftover_check.cpp"), "-o", exe], check=True, capture_output=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
    assert out.stdout.strip().endswith("0 bad")
