"""CPU test of the host-side layout builders (csrc/layout.hpp): row partition, short
tiles and the XCD column slices (balanced, contiguous, order-preserving, tiles that
respect the long/short boundary).  Compiles tests/layout_check.hip for the host only
and runs it; no GPU is needed."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not shutil.which(HIPCC) and not os.path.exists(HIPCC), reason="hipcc not found")
def test_layout_invariants(tmp_path):
    exe = tmp_path / "layout_check"
    src = os.path.join(HERE, "layout_check.hip")
    subprocess.run([HIPCC, "-O1", "-std=c++17", "--offload-arch=gfx950", src, "-o", str(exe)],
                   check=True, capture_output=True, timeout=300)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout + out.stderr
