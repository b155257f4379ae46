"""Trace order of stratified sweeps (st_common.h), checked on the CPU by
tests/cpp/order_check.cpp: the Hilbert curve over the 256 x 256 sun-plane
cells visits each cell once with unit steps, and stratified_group_of_cell
inverts the keyed Feistel cell permutation of the stratified sampler, so that
tracing a sweep in Hilbert order covers exactly its rays (the per-ray results
do not depend on the order; the GPU parity tests compare them ray by ray)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_hilbert_order_and_feistel_inverse(tmp_path):
    exe = str(tmp_path / "order_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wno-unknown-pragmas", "-I", os.path.join(ROOT, "include"), "-I",
                    os.path.join(ROOT, "optixsoltracedemos_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "order_check.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], stdout=subprocess.PIPE, text=True, check=True).stdout.split()
    r = {out[i]: int(out[i + 1]) for i in range(1, len(out), 2)}
    assert r == {"cells": 65536, "steps_not_adjacent": 0, "inverse_mismatch": 0}, r
