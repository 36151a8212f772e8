"""The C++ binding (include/talus_mi.hpp over include/talus_mi.h) compiles
with g++ against libtalus_mi.so and drives the library the way a
talus::World host would (tests/cpp/world_binding.cpp)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def program(tmp_path_factory):
    lib = os.path.join(ROOT, "talus_amd")
    out = str(tmp_path_factory.mktemp("cpp") / "world_binding")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "world_binding.cpp"), "-L" + lib, "-ltalus_mi",
           "-Wl,-rpath," + lib, "-o", out]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return out


def run(program):
    return subprocess.run([program], check=True, capture_output=True, text=True, timeout=120).stdout


def test_cpp_binding_builds_the_mesh(program):
    out = run(program)
    m = re.search(r"mesh (\d+) nodes (\d+) tets mass ([0-9.]+)", out)
    assert m and int(m.group(1)) == 275 and int(m.group(2)) == 960
    assert abs(float(m.group(3)) - 1000.0 * 0.8 * 0.2 * 0.2) < 1e-4


@pytest.mark.gpu
def test_cpp_binding_steps_on_the_device(program):
    out = run(program)
    m = re.search(r"stepped 10 steps nonfinite (\d+) sag (\S+)", out)
    assert m, out
    assert int(m.group(1)) == 0 and float(m.group(2)) < 0.0


@pytest.mark.gpu
def test_cpp_binding_finds_ground_contacts(program):
    out = run(program)
    m = re.search(r"contacts (\d+) lowest gap (\S+) lowest node (\S+)", out)
    assert m, out
    # a plane 0.3 m below the beam's bottom face (11 x 5 nodes): every bottom
    # node touches it once, the sagging tip closest, at its own height above
    # the plane (the closest point of a triangle to a plane is its lowest corner)
    n, gap, node = int(m.group(1)), float(m.group(2)), float(m.group(3))
    assert n >= 55 and 0.0 < gap <= 0.3
    assert abs(gap - node) <= 1e-4
