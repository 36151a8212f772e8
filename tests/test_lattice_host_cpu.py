"""Host side of the strainstep hex lattices on the CPU: the library's
tearCells and lowestAngular (talus_amd/csrc/hex_build.cpp) equal the hex
oracle's restatements bit for bit (tests/cpp/lattice_host_check.cpp, built
with g++ against libtalus_mi.so and the oracle)."""
import os
import subprocess

import pytest

from oracle import hex as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def program(tmp_path_factory):
    H.lib()  # builds the oracle when it is missing
    lib, orc = os.path.join(ROOT, "talus_amd"), os.path.join(ROOT, "oracle")
    if not os.path.exists(os.path.join(lib, "libtalus_mi.so")):
        pytest.skip("libtalus_mi.so is not built")
    out = str(tmp_path_factory.mktemp("cpp") / "lattice_host_check")
    cmd = ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           os.path.join(ROOT, "tests", "cpp", "lattice_host_check.cpp"), "-L" + lib, "-ltalus_mi", "-L" + orc,
           "-loracle", "-Wl,-rpath," + lib + ":" + orc, "-o", out]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return out


def test_library_tear_and_lowest_mode_equal_the_oracle(program):
    r = subprocess.run([program], capture_output=True, text=True, timeout=300)
    assert "tear different 0 of 40" in r.stdout, r.stdout[-2000:]
    assert "lowest different 0 of 6" in r.stdout, r.stdout[-2000:]
    assert r.returncode == 0
