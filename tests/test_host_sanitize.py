"""Host-side C++ (the graph builder and its file format, the CPU oracle) built with
AddressSanitizer and UndefinedBehaviorSanitizer and driven through construction, persistence,
every oracle search mode and the error paths (tests/cpp/host_sanitize_test.cpp).  No GPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.timeout(900)
def test_host_code_under_asan_and_ubsan(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "host_sanitize_test")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-ffp-contract=off", "-pthread",
           "-I" + os.path.join(ROOT, "hnsw_rabitq_amd", "csrc"), "-I" + os.path.join(ROOT, "oracle"),
           os.path.join(ROOT, "tests", "cpp", "host_sanitize_test.cpp"),
           os.path.join(ROOT, "hnsw_rabitq_amd", "csrc", "host", "hnsw_index.cpp"),
           os.path.join(ROOT, "oracle", "oracle.cpp"), "-o", exe]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout[-4000:]
    env = dict(os.environ, SANITIZE_TMP=str(tmp_path), ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:allocator_may_return_null=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env, timeout=800)
    assert "ERROR: AddressSanitizer" not in r.stdout and "runtime error" not in r.stdout, r.stdout[-6000:]
    assert r.returncode == 0 and "host sanitize: ok" in r.stdout, r.stdout[-6000:]


@pytest.mark.timeout(900)
def test_host_code_under_tsan(tmp_path):
    """The same program under ThreadSanitizer (construction with 4 threads, the oracle's threaded
    searches): no data race.  Lock-order reports are off: an inserter holds its new element's
    list lock while it takes other lists' locks for a moment (the reference's locking in
    addPoint, hnswalg.h:1152-1266), and every such pair is reported as a potential cycle."""
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "host_tsan_test")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=thread", "-ffp-contract=off",
           "-pthread", "-I" + os.path.join(ROOT, "hnsw_rabitq_amd", "csrc"), "-I" + os.path.join(ROOT, "oracle"),
           os.path.join(ROOT, "tests", "cpp", "host_sanitize_test.cpp"),
           os.path.join(ROOT, "hnsw_rabitq_amd", "csrc", "host", "hnsw_index.cpp"),
           os.path.join(ROOT, "oracle", "oracle.cpp"), "-o", exe]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout[-4000:]
    env = dict(os.environ, SANITIZE_TMP=str(tmp_path), TSAN_OPTIONS="detect_deadlocks=0:halt_on_error=0")
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env, timeout=800)
    assert "WARNING: ThreadSanitizer" not in r.stdout, r.stdout[-6000:]
    assert r.returncode == 0 and "host sanitize: ok" in r.stdout, r.stdout[-4000:]
