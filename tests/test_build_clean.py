"""The build bootstraps from a clean checkout: in a copy of the tracked files
(no built libraries), __graft_entry__.build() compiles libnpx_hip.so, the
host module, the oracle and the arithmetic check, and the package imports."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.slow
def test_build_from_clean_checkout(tmp_path):
    if shutil.which("git") is None or not os.path.isdir(os.path.join(ROOT, ".git")):
        pytest.skip("not a git checkout")
    files = subprocess.run(["git", "ls-files", "-z"], cwd=ROOT, capture_output=True, check=True).stdout
    dst = tmp_path / "repo"
    for rel in filter(None, files.decode().split("\0")):
        src = os.path.join(ROOT, rel)
        if not os.path.isfile(src):
            continue
        os.makedirs(os.path.dirname(dst / rel), exist_ok=True)
        shutil.copy2(src, dst / rel)
    built = [os.path.join(d, f) for d, _, fs in os.walk(dst) for f in fs if f.endswith(".so")]
    assert not built, built
    code = ("import __graft_entry__ as g; g.build(); import novaphy_amd as nv; "
            "print(nv.FluidWorld.__name__, nv.device_count() >= 0)")
    r = subprocess.run([sys.executable, "-c", code], cwd=dst, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert "FluidWorld True" in r.stdout
    for rel in ("novaphy_amd/libnpx_hip.so", "oracle/libpbf_oracle.so", "tests/hip/libnpx_arith_check.so"):
        assert os.path.isfile(dst / rel), rel
