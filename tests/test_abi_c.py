"""The C-ABI from C++: tests/abi_c_check.cpp includes include/mpqc_ccsd.h, links
libmpqc_ccsd.so and drives it the way a C++ host (MPQC's CCSD::compute_ccsd)
would: error paths without a GPU, and on the GPU one residual in TiledArray
tile layout and one converged solve against the oracle's values
(tests/golden/abi_residual_o2_v5.txt, written by tests/golden/make_abi_fixture.py).
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "abi_c_check.cpp")
LIBDIR = os.path.join(ROOT, "mpqc_amd", "lib")
PREBUILT = os.path.join(LIBDIR, "abi_c_check")
FIXTURE = os.path.join(ROOT, "tests", "golden", "abi_residual_o2_v5.txt")


def build_checker(out=PREBUILT):
    """g++ against the header and the built library (rpath next to it)."""
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(ROOT, "include"), SRC, "-L", LIBDIR,
           "-lmpqc_ccsd", f"-Wl,-rpath,{LIBDIR}", "-o", out]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return out


def _checker(tmp_path_factory):
    lib = os.path.join(LIBDIR, "libmpqc_ccsd.so")
    if not os.path.exists(lib):
        pytest.fail(f"{lib} is not built")
    if (os.path.exists(PREBUILT) and os.path.getmtime(PREBUILT) >= os.path.getmtime(SRC)
            and os.path.getmtime(PREBUILT) >= os.path.getmtime(lib)):
        return PREBUILT
    return build_checker(str(tmp_path_factory.mktemp("abi") / "abi_c_check"))


def _run(args):
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    print(r.stdout)
    print(r.stderr)
    return r


def test_c_abi_error_paths(tmp_path_factory):
    r = _run([_checker(tmp_path_factory), "cpu"])
    assert r.returncode == 0 and "FAIL" not in r.stdout


@pytest.mark.gpu
def test_c_abi_residual_and_solve_from_cpp(tmp_path_factory):
    r = _run([_checker(tmp_path_factory), "gpu", FIXTURE])
    assert r.returncode == 0 and "FAIL" not in r.stdout
    assert "residual matches the oracle" in r.stdout and "converged energy matches" in r.stdout
