"""The C++ drop-in classes (include/hnswgpu.hpp): the reference's searchKnnCloserFirst invariant
(tests/cpp/searchKnnCloserFirst_test.cpp:46-68) for HierarchicalNSW and BruteforceSearch, batched
and filtered searches, run by the program tests/cpp/closer_first_test.cpp (built in-tree by
hnsw_rabitq_amd/build.py)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = os.path.join(ROOT, "tests", "cpp", "closer_first_test")


def test_header_compiles():
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
                        "-x", "c++", "-"], input='#include "hnswgpu.hpp"\nint main() { return 0; }\n',
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_abi_driver_compiles():
    """tests/cpp/abi_drive.cpp (the C ABI driver of the host-sanitizer check,
    tests/cpp/host_asan_gpu.sh) stays in step with the header."""
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp", "abi_drive.cpp")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_cpp_closer_first_invariant():
    assert os.path.exists(PROG), "build the library first (hnsw_rabitq_amd/build.py builds the test program)"
    print("case=cpp closer-first invariant", flush=True)
    r = subprocess.run([PROG], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and "Test ok" in r.stdout
