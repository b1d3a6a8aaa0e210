"""The literal path's float arctangent (csrc/rsf_atan.hpp: a degree-9
polynomial in double after a one-division range reduction) against the
oracle's det_atan rounded to float, on every one of the 2^32 float inputs.
The host build of the header divides with `/`; the device build's unscaled
division sequence is the compiler's IEEE division minus steps that are
identities for its operand range, so the check covers both."""
import os
import subprocess

import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    oracle.build()
    exe = str(tmp_path_factory.mktemp("atan") / "check_atan")
    odir = os.path.join(ROOT, "oracle")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fopenmp", "-ffp-contract=off",
                    "-I", os.path.join(ROOT, "paper-rsf-vesselseg_amd", "csrc"), "-I", odir,
                    os.path.join(HERE, "check_atan.cpp"), "-o", exe, "-L", odir, "-loracle",
                    "-Wl,-rpath," + odir], check=True)
    return exe


def test_fast_atan_equals_oracle_on_every_float(checker):
    r = subprocess.run([checker, "1"], capture_output=True, text=True, timeout=600)
    checked, mism, max_ulp = r.stdout.split()
    assert int(checked) == 2 ** 32
    assert int(mism) == 0 and r.returncode == 0, r.stderr
    # the polynomial stays within a few units of the Cephes double
    assert float(max_ulp) <= 4
