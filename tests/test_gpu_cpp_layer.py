"""The hip:: C++ components driven like cuda::md::MD (tests/cpp/test_hip_md.cpp):
fused device loop and per-phase component walk give bit-identical
trajectories; energy is conserved; a missing neighbor list is rejected."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_hip_md_cpp_consumer():
    exe = os.path.join(ROOT, "tests", "cpp", "test_hip_md")
    assert os.path.exists(exe), "tests/cpp/test_hip_md not built (run __graft_entry__.build())"
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0, out.stderr
    assert "per-phase == fused" in out.stdout
    assert "per-phase == fused (PME)" in out.stdout
    assert "hip::path::Remd: 12 attempts" in out.stdout
    assert "hip::md::Virial: P " in out.stdout


def test_hip_driver_boundary():
    """Driver<md::Phase, hip::device_tag> equals hip::md::MD bit for bit;
    StopOnNaN ends a run; NoseHooverNPT gets an injected Virial and reaches
    the reference regression's volume; thermo, singular setters, memory and
    protocol accessors; run() on a cpu::Frame-like frame (tests/cpp/
    test_hip_driver.cpp)."""
    exe = os.path.join(ROOT, "tests", "cpp", "test_hip_driver")
    assert os.path.exists(exe), "tests/cpp/test_hip_driver not built (run __graft_entry__.build())"
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0, out.stderr
    assert "Driver<md::Phase, hip::device_tag> == hip::md::MD" in out.stdout
    assert "StopOnNaN ended the fused run after 38 of 100 steps" in out.stdout
    assert "StopOnNaN ended the per-phase run after 38 of 100 steps" in out.stdout
    assert "Virial injected" in out.stdout
    assert "all checks passed" in out.stdout
