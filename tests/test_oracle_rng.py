"""Oracle RNG pins: Philox4x32-10 against the published Random123 known-answer
vectors and against rocRAND's open-source implementation; the uniform mapping;
Halton radical inverses against closed forms (cuda/sun.cu:16-24)."""
import json
import os
import random
import shutil
import subprocess
from fractions import Fraction

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_philox_known_answers(oracle_mod):
    kat = json.load(open(os.path.join(GOLDEN, "philox4x32_10_kat.json")))
    for v in kat["vectors"]:
        out = oracle_mod.philox([int(x, 16) for x in v["ctr"]], [int(x, 16) for x in v["key"]])
        assert out == [int(x, 16) for x in v["out"]]


def test_philox_matches_rocrand(oracle_mod, tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available to build the rocRAND cross-check")
    exe = tmp_path / "xc"
    subprocess.run([hipcc, "-x", "hip", "--offload-arch=gfx950", "-O1",
                    os.path.join(GOLDEN, "rocrand_philox_xcheck.cpp"), "-o", str(exe)],
                   check=True, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    rng = random.Random(7)
    cases = [[rng.getrandbits(32) for _ in range(6)] for _ in range(200)]
    # the counter layout the tracer uses: (ray_lo, ray_hi, block, job)
    cases += [[i, 0, b, j, 0, 0] for i in (0, 1, 2**31, 2**32 - 1) for b in (0, 1, 2, 9) for j in (0, 5)]
    inp = "\n".join(" ".join(f"{w:08x}" for w in c) for c in cases) + "\n"
    out = subprocess.run([str(exe)], input=inp, stdout=subprocess.PIPE, text=True, check=True).stdout.split("\n")
    for c, line in zip(cases, out):
        ref = [int(x, 16) for x in line.split()]
        assert oracle_mod.philox(c[:4], c[4:]) == ref


def test_uniform_mapping_exact(oracle_mod):
    for x in [0, 1, 511, 512, 2**31, 2**32 - 512, 2**32 - 1, 0xDEADBEEF]:
        u = oracle_mod.uniform(x)
        assert 0.0 < u < 1.0
        k = u * 2**24
        assert k == int(k) and int(k) % 2 == 1          # (2j+1)/2^24: exact with 24 mantissa bits
        assert np.float32(u) == u                       # exactly representable in FP32
        assert np.float32(np.float32(1.0) - np.float32(u)) == 1.0 - u
    assert oracle_mod.uniform(0) == 0.5 / 2**23
    assert oracle_mod.uniform(2**32 - 1) == 1.0 - 0.5 / 2**23


def _radical_inverse(i, b):
    f, r = Fraction(1, 1), Fraction(0, 1)
    while i > 0:
        f /= b
        r += f * (i % b)
        i //= b
    return r


def test_halton_closed_forms(oracle_mod):
    assert oracle_mod.halton(0, 2) == 0.0 and oracle_mod.halton(0, 3) == 0.0
    assert oracle_mod.halton(1, 2) == 0.5
    assert oracle_mod.halton(2, 2) == 0.25
    assert oracle_mod.halton(3, 2) == 0.75
    assert oracle_mod.halton(1, 3) == pytest.approx(1 / 3, abs=1e-16)
    assert oracle_mod.halton(2, 3) == pytest.approx(2 / 3, abs=1e-16)
    assert oracle_mod.halton(3, 3) == pytest.approx(1 / 9, abs=1e-16)
    for i in [5, 17, 1000, 123456, 2**31 - 1, 2**32 - 1]:
        for b in (2, 3):
            assert oracle_mod.halton(i, b) == float(_radical_inverse(i, b))
    # base 2 is the bit reversal of the index over 2^32
    for i in [1, 6, 12345, 2**32 - 1]:
        rev = int(f"{i:032b}"[::-1], 2)
        assert oracle_mod.halton(i, 2) == rev / 2**32
