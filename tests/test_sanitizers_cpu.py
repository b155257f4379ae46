"""Host code under AddressSanitizer and UBSan (gcc/g++ only; no device code):
the FP64 oracle (oracle/oracle.c) tracing the three-heliostat scene, the dish
and a 2,000-heliostat field with records and every planted mistake, the BVH
topology builder (st_bvh.cpp) over random and degenerate box sets, and the
C++ drop-in layer's STINPUT reader over the fixture and malformed files.
Any sanitizer report aborts the program and fails the test."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "optixsoltracedemos_amd")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")


def libasan():
    p = subprocess.run(["gcc", "-print-file-name=libasan.so"], stdout=subprocess.PIPE, text=True).stdout.strip()
    if not os.path.isabs(p) or not os.path.exists(p):
        pytest.skip("gcc has no AddressSanitizer runtime here")
    return p


def test_oracle_under_sanitizers(tmp_path):
    so = str(tmp_path / "liboracle_san.so")
    subprocess.run(["gcc", "-std=c11", "-fPIC", "-shared", *SAN, os.path.join(ROOT, "oracle", "oracle.c"),
                    "-o", so, "-lm", "-lpthread"], check=True)
    script = r"""
import numpy as np
from optixsoltracedemos_amd import scenes
from oracle import oracle as O
assert O.LIB.endswith("liboracle_san.so")
for w in (scenes.three_heliostats(20_000), scenes.parabolic_dish(20_000), scenes.heliostat_field(2000, 20_000)):
    s = O.Scene(w.elements)
    for pos in (0, 1, 2):
        s.set_sun(w.sun_vector, w.sun_param, w.sun_shape, position_sampling=pos)
        r = s.trace(w.num_rays, max_depth=w.max_depth, flux_nx=w.flux_nx, flux_ny=w.flux_ny, records=True, threads=4)
        assert r["stats"][0] == w.num_rays
    for plant in range(1, 7):
        with O.planted(plant):
            s.trace(w.num_rays, max_depth=w.max_depth, flux_nx=w.flux_nx, flux_ny=w.flux_ny, threads=2)
    s.closest_hit(w.elements[0]["origin"] + np.array([0, 0, 50.0]), [0, 0, -1.0], brute_force=True)
print("oracle ok")
"""
    env = dict(ENV, LD_PRELOAD=libasan(), ST_ORACLE_LIB=so, PYTHONPATH=ROOT)
    r = subprocess.run([sys.executable, "-c", script], env=env, cwd=ROOT, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=600)
    assert r.returncode == 0 and "oracle ok" in r.stdout, r.stdout[-4000:]


def test_bvh_builder_under_sanitizers(tmp_path):
    exe = str(tmp_path / "bvh_check")
    subprocess.run(["g++", "-std=c++17", *SAN, "-I", os.path.join(ROOT, "include"), "-I", os.path.join(PKG, "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "bvh_check.cpp"), os.path.join(PKG, "csrc", "st_bvh.cpp"),
                    "-o", exe], check=True)
    r = subprocess.run([exe], env=ENV, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout[-4000:]


def test_stinput_reader_under_sanitizers(tmp_path):
    lib = os.path.join(PKG, "_lib")
    if not os.path.exists(os.path.join(lib, "libsoltrace_hip.so")):
        pytest.fail("HIP library not built (run __graft_entry__.build())")
    exe = str(tmp_path / "dropin_parse_san")
    srcs = [os.path.join(PKG, "lib", f) for f in sorted(os.listdir(os.path.join(PKG, "lib"))) if f.endswith(".cpp")]
    subprocess.run(["g++", "-std=c++17", *SAN, "-I", os.path.join(ROOT, "include"), "-I", PKG,
                    os.path.join(ROOT, "tests", "cpp", "dropin_parse.cpp"), *srcs, "-L", lib, "-lsoltrace_hip",
                    f"-Wl,-rpath,{lib}", "-o", exe], check=True)
    from tests.test_dropin_cpu import FIXTURE, SYNTH
    synth = tmp_path / "synth.stinput"
    synth.write_text(SYNTH)
    lines = SYNTH.splitlines()
    cases = [(FIXTURE, 0), (FIXTURE, 0, "optics"), (str(synth), 0, "optics")]
    for k, cut in ((8, 10), (15, 20), (1, 1)):                  # truncated records
        bad = list(lines)
        bad[k] = "\t".join(bad[k].split("\t")[:cut])
        p = tmp_path / f"bad{k}.stinput"
        p.write_text("\n".join(bad) + "\n")
        cases.append((str(p), 1))
    trunc = tmp_path / "trunc.stinput"
    trunc.write_text("\n".join(lines[:16]) + "\n")              # file ends inside a stage
    cases.append((str(trunc), 1))
    for path, want, *extra in cases:
        r = subprocess.run([exe, path, *extra], env=ENV, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                           timeout=120)
        assert r.returncode == want, (path, r.stdout[-4000:])
        assert "ERROR: AddressSanitizer" not in r.stdout and "runtime error" not in r.stdout, r.stdout[-4000:]
