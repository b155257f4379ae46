"""The atan2 of the cylindrical flux bin (st::atan2_poly in st_common.h: octant
reduction and the polynomial of Abramowitz & Stegun eq. 4.4.49) on the host:
its largest error against libm's atan2 over 1.2e7 FP32 inputs stays at the
FP32 rounding of a result near pi (one ulp of pi is 2.4e-7 rad), the axes and
signed zeros come out exact, and a planted mistake (the last coefficient
dropped) is caught."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("atan") / "atan_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"), "-I",
                    os.path.join(ROOT, "optixsoltracedemos_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "atan_check.cpp"), "-o", out], check=True)
    return out


def run(exe, *args):
    t = subprocess.run([exe, *args], stdout=subprocess.PIPE, text=True, check=True).stdout.split()
    return {t[i]: float(t[i + 1]) for i in range(0, len(t) - 1, 2) if t[i] != "at_theta"}


def test_atan2_poly_error_bound(exe):
    r = run(exe)
    assert r["samples"] > 1e7
    assert r["max_error_rad"] < 4e-7
    assert r["axis_case_failures"] == 0


def test_atan2_check_catches_a_dropped_coefficient(exe):
    r = run(exe, "planted")
    assert r["max_error_rad"] > 1e-4 and r["axis_case_failures"] > 0
