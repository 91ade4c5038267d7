"""The driver I/O layer (include/ed/driver_io.hpp and its Python mirror driver_io.py).

* ed::argsparse against the reference's own MBL_test/include/argparse.h, compiled into
  oracle/_ref/ref_argparse (g++, host only): both run the TotalSpin drivers' option set on the
  same argument lists and must print the same values, the same errors and exit the same way.
* randNum files: our writer's format against the reference's generator (oracle/_ref/
  ref_randnum_generator, which seeds from std::random_device, so only the format is compared);
  the C++ and Python writers byte for byte for a seed; the values against std::mt19937's
  known-answer output (10000th draw of the default seed 5489 is 4123659995).
* replay: the header line is the first sample (zeros), then the file's rows, then zeros.
* .dat files: header once, rows appended, C++ and Python byte for byte.
"""
import os
import subprocess

import numpy as np
import pytest

from exact_diagonalization_for_fermionic_systems_with_gpu_amd import driver_io as DIO

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_AP = os.path.join(REPO, "oracle", "_ref", "ref_argparse")
REF_GEN = os.path.join(REPO, "oracle", "_ref", "ref_randnum_generator")


@pytest.fixture(scope="module")
def check_exe(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("dio") / "driver_io_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(REPO, "include"), "-o", exe,
                    os.path.join(REPO, "tests", "cpp_host", "driver_io_check.cpp")], check=True)
    return exe


def _ref(path):
    if not os.path.exists(path) and os.path.isdir("/root/reference"):
        os.system(f"make -s -C {os.path.join(REPO, 'oracle')} ref >/dev/null 2>&1")
    if not os.path.exists(path):
        pytest.skip("reference host programs (oracle/_ref) not built")
    return path


CASES = [
    ["-U=4", "-Q=6,6", "-dev=0"],
    ["-U=4.5", "-Q=5,7", "-dev=1", "-X=0.25", "-Nsam=3", "-path=/tmp/out", "-xpath=/tmp/x"],
    ["-Q=6,6", "-dev=0", "-U=8", "-X=1e-3"],
    ["-U=4", "-Q=6,6"],                      # missing option without default
    ["-U=4", "-Q=6,6", "-dev=0", "-U=5"],    # repeated option
    ["-U4", "-Q=6,6", "-dev=0"],             # no '='
    ["-U=", "-Q=6,6", "-dev=0"],             # empty value
    ["-U=4", "-Q=6,6,", "-dev=0"],           # trailing comma in a list
    ["-U=4", "-Q=6", "-dev=0"],              # mfind<int, 2> of a single value
    ["-U=abc", "-Q=6,6", "-dev=0"],          # lexical cast failure
    ["-U=4", "-Q=6,6", "-dev=0", "-Xpath=3"],  # 'X' claims the argument, then no '='
    ["-U=4", "-Q=6,6", "-dev=0", "--help"],
    ["+U=4", "-Q=6,6", "-dev=0", "-unknown=1"],  # any leading character; unknown names ignored
]


@pytest.mark.parametrize("k", range(len(CASES)))
def test_argsparse_matches_reference(check_exe, k):
    ref = _ref(REF_AP)
    args = CASES[k]
    r = subprocess.run([ref] + args, capture_output=True, text=True)
    o = subprocess.run([check_exe, "parse"] + args, capture_output=True, text=True)
    norm = lambda s: s.replace(ref, "PROG").replace(check_exe, "PROG")
    assert o.returncode == r.returncode
    assert norm(o.stdout) == norm(r.stdout)
    assert norm(o.stderr) == norm(r.stderr)


def test_randnum_format_matches_reference_generator(check_exe, tmp_path):
    gen = _ref(REF_GEN)
    subprocess.run([gen, "-L=12", "-Nsam=3", f"-path={tmp_path}"], check=True)
    ref = open(tmp_path / "randNum_Kagome-L12.dat").read().split("\n")
    ours = tmp_path / "ours.dat"
    subprocess.run([check_exe, "randnum", str(ours), "12", "3", "7"], check=True)
    mine = open(ours).read().split("\n")
    assert ref[0] == mine[0] == DIO.RANDNUM_HEADER
    assert len(ref) == len(mine) == 5 and ref[-1] == mine[-1] == ""
    for a, b in zip(ref[1:4], mine[1:4]):
        fa, fb = a.split("\t"), b.split("\t")
        assert len(fa) == len(fb) == 13 and fa[-1] == fb[-1] == ""  # 12 tab-terminated numbers
        for t in fa[:-1] + fb[:-1]:
            # printed as operator<< with precision 10 prints: the %.10g text of the value
            assert -1.0 <= float(t) < 1.0 and format(float(t), ".10g") == t, t


def test_randnum_cpp_and_python_identical_and_append(check_exe, tmp_path):
    a, b = tmp_path / "a.dat", tmp_path / "b.dat"
    for _ in range(2):  # the second call appends rows without a second header
        subprocess.run([check_exe, "randnum", str(a), "14", "4", "12345"], check=True)
        DIO.randnum_write(str(b), 14, 4, 12345)
    ta, tb = open(a).read(), open(b).read()
    assert ta == tb
    assert ta.count("#") == 1 and ta.count("\n") == 9


def test_mt19937_known_answer():
    # std::mt19937 with the default seed 5489: the 10000th output is 4123659995 (C++11 26.5.5)
    rs = np.random.RandomState(5489)
    assert int(rs.randint(0, 2 ** 32, size=10000, dtype=np.uint64)[-1]) == 4123659995
    x = DIO.mt19937_uniform(5489, 1000)
    assert np.all((x >= -1.0) & (x < 1.0))


def test_replay_reads_header_as_first_sample(check_exe, tmp_path):
    f = tmp_path / "r.dat"
    DIO.randnum_write(str(f), 5, 2, 99)
    rows = [[float(t) for t in line.split("\t")[:-1]] for line in open(f).read().split("\n")[1:3]]
    out = subprocess.run([check_exe, "replay", str(f), "5", "4"], capture_output=True, text=True, check=True)
    got = [[float(t) for t in line.split()] for line in out.stdout.strip().split("\n")]
    assert got[0] == [0.0] * 5
    assert got[1] == rows[0] and got[2] == rows[1]
    assert got[3] == [0.0] * 5
    rp = DIO.RandNumReplay(str(f))
    assert [list(rp.next(5)) for _ in range(4)] == got
    miss = subprocess.run([check_exe, "replay", str(tmp_path / "none.dat"), "3", "1"], capture_output=True,
                          text=True, check=True)
    assert miss.stdout.startswith("There is not the file!")


def test_dat_file_cpp_and_python_identical(check_exe, tmp_path):
    a, b = tmp_path / "a.dat", tmp_path / "b.dat"
    for _ in range(2):
        subprocess.run([check_exe, "dat", str(a)], check=True)
        d = DIO.DatFile(str(b), "#       E       (S^+S^- + S^-S^+)/2       S(S+1)    ")
        d.row([-8.123456789012345, 1.5e-12, 2.0])
        d.row([-7.0, 0.75, 0.75])
        d.close()
    assert open(a).read() == open(b).read()
    lines = open(a).read().split("\n")
    assert lines[1] == "\t-8.123456789\t1.5e-12\t2" and len(lines) == 6
