"""The 2-D Roe flux of the stage kernel (ox_physics.hpp roe_local2) returns,
for every face state, the values of the 3-D routine roe_local with the third
velocity component zero, with the tangential component in the t1 slot (x
faces) and in the t2 slot (y faces), in FP64 and FP32: the host build of the
same device functions on 300,000 random states (including equal left and
right states and zero tangential velocity)."""
import pathlib
import shutil
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_roe_local2_equals_roe_local_with_a_zero_component(tmp_path):
    exe = tmp_path / "roe2_check"
    cmd = ["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
           f"-I{ROOT / 'ouxsbli_amd' / 'csrc'}", f"-I{ROOT / 'include'}", "-x", "c++",
           str(ROOT / "tests" / "cpp" / "roe2_check.cpp"), "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "bad=0 badf=0" in r.stdout
