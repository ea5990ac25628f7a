"""The device octree build's algorithm on the CPU (tests/octree_keys_check.cpp):
descent keys, the tree's shape decided from the sorted keys' counts, and the
lists gathered from per-level sorts, against the host build (itself pinned to
the oracle by tests/test_octree.py).  The same checker's device half runs in
tests/test_gpu_tree.py."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "beep3_amd" / "csrc"


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_device_octree_algorithm_matches_host_build(tmp_path):
    exe = tmp_path / "octree_keys_check"
    cmd = ["g++", "-O2", "-std=c++17", "-pthread", f"-I{CSRC}",
           str(ROOT / "tests" / "octree_keys_check.cpp"), str(CSRC / "host_octree.cpp"),
           str(CSRC / "host_mesh.cpp"), "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=300)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("identical") == 5
