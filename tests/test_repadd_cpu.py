"""npx_rep_add (novaphy_amd/csrc/npx_repadd.hpp), the exact k-fold float
addition the neighbour passes use for runs of identical candidates, compiled
for the host with g++ and checked bit for bit against k sequential additions
(tests/cpp/repadd_check.cpp): random magnitudes and signs, exact ties, zero
and binade crossings, subnormals, inf and NaN."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_rep_add_equals_sequential_additions(tmp_path):
    exe = str(tmp_path / "repadd_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-o", exe,
                    os.path.join(HERE, "cpp", "repadd_check.cpp")], check=True)
    r = subprocess.run([exe, "100000"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:]
    assert " 0 mismatches" in r.stdout
