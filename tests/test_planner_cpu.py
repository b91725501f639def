"""The cache-blocked pass planner on the host (no device): tests/plan_emulator.cpp
runs planned passes with the kernels' op semantics against the oracle over
random circuits of the whole gate table, tile sizes 3..12 and both register
segment widths."""
import glob
import os
import subprocess

import pytest

from conftest import ROOT

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def test_planned_passes_reproduce_circuits(tmp_path):
    objs = sorted(glob.glob(os.path.join(ROOT, "moonlab_amd", "_build", "*.o")))
    if not objs or not os.path.exists(HIPCC):
        pytest.skip("engine objects not built (python -m moonlab_amd.build)")
    inc = ["-I" + os.path.join(ROOT, d) for d in ("include", "moonlab_amd/csrc", "oracle")]
    exe = str(tmp_path / "plan_emulator")
    subprocess.run(["gcc", "-O2", "-march=x86-64-v3", "-fopenmp", "-c", os.path.join(ROOT, "oracle", "sv_oracle.c"),
                    "-I" + os.path.join(ROOT, "oracle"), "-o", str(tmp_path / "svo.o")], check=True)
    subprocess.run([HIPCC, "-O2", "-std=c++17", *inc, "-c", os.path.join(ROOT, "tests", "plan_emulator.cpp"),
                    "-o", str(tmp_path / "pe.o")], check=True, capture_output=True)
    subprocess.run([HIPCC, "--offload-arch=gfx950", str(tmp_path / "pe.o"), str(tmp_path / "svo.o"), "-x", "none",
                    *objs, "-o", exe, "-lpthread", "-fopenmp"], check=True, capture_output=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok:"), r.stdout
    # the register-phase programs include paired rotations (ARM_ROT2)
    assert int(r.stdout.split(",")[-1].split()[0]) > 0, r.stdout
