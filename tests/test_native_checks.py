"""Host-side checks of device arithmetic that the kernels share with the
host compiler: built with hipcc for the host, run on the CPU (no GPU)."""
import shutil
import subprocess
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parents[1]
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not Path(HIPCC).exists(), reason="hipcc not installed")
def test_batch_radius_bound_never_misnames_a_point(tmp_path):
    """batch_const_radius() only hoists a radius that every decodable point of
    the batch computes (random cameras, sizes, model matrices, boxes)."""
    exe = tmp_path / "radius_bound_check"
    src = REPO / "tests" / "native" / "radius_bound_check.cpp"
    subprocess.run([HIPCC, "-x", "hip", "--offload-arch=gfx950", "-O2", "-std=c++17", "-ffp-contract=off",
                    "-fno-fast-math", "-I", str(REPO / "include"), "-I", str(REPO / "satin-computerasteriser_amd" / "csrc"),
                    str(src), "-o", str(exe)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "hoisted" in out.stdout


@pytest.mark.skipif(not Path(HIPCC).exists(), reason="hipcc not installed")
def test_batch_bounds_hold_for_every_point(tmp_path):
    """batch_bounds() (the cull pass's proof of a batch's fragment rectangle,
    depth bound, clip/viewport pass and reciprocal range) against the draw
    passes' exact per-point arithmetic, on ordinary and far-offset/close-camera
    batches; without its rounding allowance the same cases must fail."""
    exe = tmp_path / "batch_bounds_check"
    src = REPO / "tests" / "native" / "batch_bounds_check.cpp"
    subprocess.run([HIPCC, "-x", "hip", "--offload-arch=gfx950", "-O2", "-std=c++17", "-ffp-contract=off",
                    "-fno-fast-math", "-I", str(REPO / "include"), "-I", str(REPO / "satin-computerasteriser_amd" / "csrc"),
                    str(src), "-o", str(exe)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "bounds hold" in out.stdout
