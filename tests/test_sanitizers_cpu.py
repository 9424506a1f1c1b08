"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the CPU code: the
oracle (oracle/*.cpp), the C++ host's CPU paths (rigid pipeline, sweep and
prune, narrowphase, samplers, performance monitor) and npx_rep_add, each
driven by a small program (tests/cpp/*_sanitize.cpp, repadd_check.cpp).  A
sanitizer report fails the run (halt_on_error / -fno-sanitize-recover)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="halt_on_error=1:detect_leaks=1", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")

pytestmark = [pytest.mark.slow, pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")]


def _run(exe, *args):
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=600, env=ENV)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-6000:])
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-6000:]
    return r.stdout


def test_oracle_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "oracle_sanitize")
    subprocess.run(["g++", "-std=c++17", "-ffp-contract=off", *SAN, "-o", exe,
                    os.path.join(HERE, "cpp", "oracle_sanitize.cpp"), os.path.join(ROOT, "oracle", "pbf_oracle.cpp"),
                    os.path.join(ROOT, "oracle", "rigid_oracle.cpp")], check=True)
    assert "oracle driver ok" in _run(exe)


def test_host_cpu_paths_under_asan_ubsan(tmp_path):
    lib = os.path.join(ROOT, "novaphy_amd", "libnpx_hip.so")
    if not os.path.exists(lib):
        pytest.skip("libnpx_hip.so is not built")
    host = os.path.join(ROOT, "novaphy_amd", "csrc", "host")
    srcs = [os.path.join(host, f) for f in ("rigid.cpp", "fluid.cpp", "perf.cpp")]
    exe = str(tmp_path / "host_sanitize")
    subprocess.run(["g++", "-std=c++20", *SAN, "-I", os.path.join(ROOT, "include"), "-I", host, "-o", exe,
                    os.path.join(HERE, "cpp", "host_sanitize.cpp"), *srcs, "-L", os.path.dirname(lib),
                    "-l:libnpx_hip.so", "-Wl,-rpath," + os.path.dirname(lib)], check=True)
    assert "host driver ok" in _run(exe, str(tmp_path / "trace.json"))


def test_repadd_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "repadd_sanitize")
    subprocess.run(["g++", "-std=c++17", "-ffp-contract=off", *SAN, "-o", exe,
                    os.path.join(HERE, "cpp", "repadd_check.cpp")], check=True)
    assert " 0 mismatches" in _run(exe, "20000")
