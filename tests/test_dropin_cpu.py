"""C++ drop-in layer on the CPU: the STINPUT reader (format of
data/stinput/toy_problem_parabolic.stinput, parser lib/soltrace_system.cpp:449-846),
element frames against the oracle, reference AABBs and GeometryDataST types."""
import os
import subprocess

import numpy as np
import pytest

from optixsoltracedemos_amd import _abi, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "optixsoltracedemos_amd")
FIXTURE = os.path.join(ROOT, "tests", "golden", "toy_problem_parabolic.stinput")


@pytest.fixture(scope="module")
def parse_exe(tmp_path_factory):
    lib = os.path.join(PKG, "_lib")
    if not os.path.exists(os.path.join(lib, "libsoltrace_system.so")):
        pytest.fail("C++ drop-in layer not built (run __graft_entry__.build())")
    exe = str(tmp_path_factory.mktemp("dropin") / "dropin_parse")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), "-I", PKG,
                    os.path.join(ROOT, "tests", "cpp", "dropin_parse.cpp"), "-L", lib, "-lsoltrace_system",
                    "-lsoltrace_hip", f"-Wl,-rpath,{lib}", "-o", exe], check=True)
    return exe


def run_parse(exe, path, optics=False):
    args = [exe, path] + (["optics"] if optics else [])
    out = subprocess.run(args, stdout=subprocess.PIPE, text=True, check=True).stdout.splitlines()
    head = out[0].split()
    elems = []
    for line in out[1:]:
        t = line.split()
        v = lambda key, n: [float(x) for x in t[t.index(key) + 1:t.index(key) + 1 + n]]
        elems.append({"surf": int(t[t.index("surf") + 1]), "o": v("o", 3), "aim": v("aim", 3), "zrot": v("zrot", 1)[0],
                      "dims": v("dims", 2), "curv": v("curv", 2), "z": v("z", 3), "x": v("x", 3),
                      "lo": v("lo", 3), "hi": v("hi", 3), "gtype": int(t[t.index("gtype") + 1]),
                      "role": int(t[t.index("role") + 1]), "rho": v("rho", 1)[0], "slope": v("slope", 1)[0],
                      "spec": v("spec", 1)[0]})
    return head, elems


def test_toy_stinput_parse(parse_exe):
    head, el = run_parse(parse_exe, FIXTURE)
    assert [float(x) for x in head[1:4]] == [0.0, 0.0, 100.0]
    assert float(head[5]) == pytest.approx(4.65e-3)                      # SIGMA (mrad) -> half-angle (rad)
    assert int(head[7]) == 0 and int(head[9]) == 3                       # pillbox, 3 heliostats
    assert len(el) == 4
    assert [e["surf"] for e in el] == [1, 1, 1, 0]                        # 3 parabolic + flat receiver
    assert el[0]["o"] == [-5.0, 0.0, 0.0] and el[0]["aim"] == pytest.approx([17.36068, -0.0, 94.72136])
    assert el[0]["zrot"] == pytest.approx(-88.757211)
    assert el[0]["dims"] == [1.0, 1.95] and el[0]["curv"] == pytest.approx([0.017068, 0.037068])
    assert el[3]["o"] == [0.0, 0.0, 10.0] and el[3]["dims"] == [2.0, 2.0]
    assert [e["gtype"] for e in el] == [2, 2, 2, 4]                       # RECTANGLE_PARABOLIC, RECTANGLE_FLAT


def test_frames_and_boxes_match_oracle(parse_exe, oracle_mod):
    _, el = run_parse(parse_exe, FIXTURE)
    for e in el:
        arr = scenes.make_elements(1)
        arr[0]["origin"], arr[0]["aim_point"], arr[0]["zrot_deg"] = e["o"], e["aim"], e["zrot"]
        arr[0]["dim_x"], arr[0]["dim_y"] = e["dims"]
        arr[0]["surface"] = _abi.SURFACE_PARABOLIC if e["surf"] == 1 else _abi.SURFACE_FLAT
        M = oracle_mod.element_frame(arr[0])
        assert np.allclose(M[:, 2], e["z"], atol=1e-11) and np.allclose(M[:, 0], e["x"], atol=1e-11)
        lo, hi = oracle_mod.element_ref_aabb(arr[0])
        assert np.allclose(lo, e["lo"], atol=1e-7) and np.allclose(hi, e["hi"], atol=1e-7)


def test_malformed_stinput_rejected(parse_exe, tmp_path):
    bad = tmp_path / "bad.stinput"
    lines = open(FIXTURE).read().splitlines()
    lines[15] = "\t".join(lines[15].split("\t")[:20])                    # element row with too few tokens
    bad.write_text("\n".join(lines) + "\n")
    r = subprocess.run([parse_exe, str(bad)], stdout=subprocess.PIPE, text=True)
    assert r.returncode == 1 and "parse failed" in r.stdout
    nohash = tmp_path / "nohash.stinput"
    nohash.write_text("\n".join(open(FIXTURE).read().splitlines()[1:]) + "\n")
    r = subprocess.run([parse_exe, str(nohash)], stdout=subprocess.PIPE, text=True)
    assert r.returncode == 1


def test_toy_stinput_optics_reach_the_element_records(parse_exe):
    """The optics the file names per element (OPTICAL rows: reflectivity column 5,
    RMS slope 7 and RMS specularity 8 in mrad, front side) reach the element
    records with set_use_stinput_optics(true); without it the elements keep
    their defaults, as the reference (which drops them, lib/soltrace_system.cpp:576-590)."""
    head, el = run_parse(parse_exe, FIXTURE, optics=True)
    assert int(head[7]) == 0                                  # SHAPE p -> pillbox
    for e in el[:3]:                                          # "Reflector"
        assert e["role"] == 0
        assert (e["rho"], e["slope"], e["spec"]) == pytest.approx((1.0, 0.95e-3, 0.2e-3))
    assert el[3]["role"] == 1 and el[3]["rho"] == 0.0         # "Absorber"
    _, plain = run_parse(parse_exe, FIXTURE)
    for e in plain:
        assert (e["rho"], e["slope"], e["spec"]) == (1.0, 0.0, 0.0)


SYNTH = """# SOLTRACE VERSION 2012.7.6 INPUT FILE
SUN\tPTSRC\t0\tSHAPE\tg\tSIGMA\t2.730000\tHALFWIDTH\t0.000000
XYZ\t0.5\t-0.25\t0.8\tUSELDH\t0\tLDH\t0.000000\t0.000000\t0.000000
USER SHAPE DATA\t2
0.0 1.0
4.65 0.5
OPTICS LIST COUNT\t1
OPTICAL PAIR\tMirror
OPTICAL\tg\t3\t1\t4\t0.930000\t0.000000\t1.500000\t0.300000\t0\t0\t0\t0\t0\t0\t1\t2
10.0 0.9
80.0 0.8
OPTICAL\tg\t3\t1\t4\t0.500000\t0.000000\t0.000000\t0.000000\t0\t0\t0\t0\t0\t0
STAGE LIST COUNT\t1
STAGE\tXYZ\t0\t0\t0\tAIM\t0\t0\t1\tZROT\t0\tVIRTUAL\t0\tMULTIHIT\t1\tELEMENTS\t3\tTRACETHROUGH\t0
Field
1\t100\t0\t0\t0\t0\t150\t0\tr\t10\t10\t0\t0\t0\t0\t0\t0\tf\t0\t0\t0\t0\t0\t0\t0\t0\t\tMirror\t2\t
1\t0\t0\t160\t0\t0\t0\t0\tc\t16\t0\t0\t0\t0\t0\t0\t0\tf\t0\t0\t0\t0\t0\t0\t0\t0\t\tMirror\t2\t
1\t0\t-8\t150\t0\t-9\t150\t0\tl\t0\t0\t20\t0\t0\t0\t0\t0\tt\t0.125\t0\t0\t0\t0\t0\t0\t0\t\tMirror\t2\t
"""


def test_stinput_cylinder_rows_cap_rows_and_tables(parse_exe, tmp_path):
    """A cylinder row ('l' section, 't' surface: radius = 1 / surface parameter 0,
    length = aperture parameter 2, centred one radius further along y), an end
    cap row ('c' + 'f', skipped), sun and optic tables (read past), the sun
    vector from XYZ and SIGMA as the half-angle."""
    f = tmp_path / "synth.stinput"
    f.write_text(SYNTH)
    head, el = run_parse(parse_exe, str(f), optics=True)
    assert [float(x) for x in head[1:4]] == [0.5, -0.25, 0.8]
    assert float(head[5]) == pytest.approx(2.73e-3) and int(head[7]) == 1      # Gaussian with optics on
    assert len(el) == 2 and int(head[9]) == 1
    m, c = el
    assert m["surf"] == 0 and m["dims"] == [10.0, 10.0] and m["o"] == [100.0, 0.0, 0.0]
    assert (m["rho"], m["slope"], m["spec"]) == pytest.approx((0.93, 1.5e-3, 0.3e-3))
    assert c["surf"] == 3 and c["role"] == 1
    assert c["o"] == [0.0, 0.0, 150.0] and c["dims"] == [16.0, 20.0]


@pytest.mark.parametrize("line,edit,msg", [
    (9, lambda t: t[:10], "line 9: OPTICAL record has 10 fields"),
    (16, lambda t: t[:8] + ["h"] + t[9:], "line 16: aperture code 'h'"),
    (16, lambda t: t[:17] + ["z"] + t[18:], "line 16: surface code 'z'"),
    (18, lambda t: t[:2] + ["north"] + t[3:], "line 18: origin is not a number"),
])
def test_stinput_errors_name_the_line(parse_exe, tmp_path, line, edit, msg):
    lines = SYNTH.splitlines()
    lines[line - 1] = "\t".join(edit(lines[line - 1].split("\t")))
    f = tmp_path / "bad.stinput"
    f.write_text("\n".join(lines) + "\n")
    r = subprocess.run([parse_exe, str(f)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    assert r.returncode == 1 and "parse failed" in r.stdout
    assert msg in r.stderr, r.stderr
