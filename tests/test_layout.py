"""CPU check of the spherical engine's flux-word layout (sph_common.hpp:
ffidx, ffown): bounds, one slot per formed face, and every read of the face
and zone updates on a written slot (tests/layout_check.cpp, built with g++)."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "black-hole-simulator_amd" / "csrc"


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_flux_layout_bounds_and_coverage(tmp_path):
    exe = tmp_path / "layout_check"
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{CSRC}", "-o", str(exe),
                    str(ROOT / "tests" / "layout_check.cpp")], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "errors 0" in r.stdout
