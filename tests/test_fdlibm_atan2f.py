"""The exact kernels' atan2 (feriphys_amd/csrc/fdlibm_atan2f.hpp) rounds like the host C
library's atan2f, which the reference's f32::atan2 calls: zero bit mismatches on special values
and on 20M random inputs (a third of them arbitrary bit patterns), for both the branch-free form
the kernels use and fdlibm's own control flow; and the division-free exact helpers equal the
division on every finite input."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_fdlibm_atan2f_matches_libm(tmp_path):
    exe = str(tmp_path / "atan2f_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                    os.path.join(HERE, "native", "atan2f_check.cpp"), "-o", exe, "-lm"], check=True)
    out = subprocess.run([exe, "20000000"], check=True, capture_output=True, text=True).stdout.strip().splitlines()
    assert out[-1] == "0", "\n".join(out[-20:])


def test_div3_rn_exhaustive(tmp_path):
    """div3_rn (the exact kernels' x / 3.0f without a division) equals x / 3.0f on all 2^32 - 2^24
    finite f32 inputs."""
    exe = str(tmp_path / "div3_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-mfma", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
                    os.path.join(HERE, "native", "div3_check.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.strip()
    assert out == "0"


def test_div_pi_rn_exhaustive(tmp_path):
    """div_pi_rn (the corner weight angle / PI without a division) equals x / PI on all finite f32
    inputs."""
    exe = str(tmp_path / "div_pi_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-mfma", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
                    os.path.join(HERE, "native", "div_pi_check.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.strip()
    assert out == "0"
