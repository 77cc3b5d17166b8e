"""Launch-workspace ownership (csrc/hcx_wspool.hpp) on the CPU: the pool's
rules run against a stand-in device that records allocations, frees and
waits (tests/cpp/test_wspool.cpp, built here with g++)."""
import os
import subprocess

from conftest import PKG, ROOT


def test_workspace_ownership_rules(tmp_path):
    exe = tmp_path / "test_wspool"
    subprocess.run(["g++", "-std=c++17", "-O1", "-pthread", "-Wall", "-I", os.path.join(PKG, "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "test_wspool.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all workspace ownership checks passed" in r.stdout
