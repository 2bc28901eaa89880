"""Host eigensolver of the EOM Davidson subspace (mpqc_amd/csrc/small_eig.hpp)
against numpy.linalg.eig: eigenvalues of random, near-diagonal and defective-free
non-symmetric matrices, and A x = lambda x for every returned eigenvector."""
import shutil
import subprocess

import numpy as np
import pytest

HERE = __import__("pathlib").Path(__file__).resolve().parent


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path_factory.mktemp("small_eig") / "small_eig_check"
    subprocess.run([gxx, "-O2", "-std=c++17", str(HERE / "small_eig_check.cpp"), "-o", str(exe)], check=True)
    return exe


def _run(exe, A):
    n = A.shape[0]
    inp = f"{n}\n" + " ".join("%.17g" % x for x in A.ravel())
    out = subprocess.run([str(exe)], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    w = np.array([complex(*map(float, line.split())) for line in out[:n]])
    X = []
    for line in out[n:2 * n]:
        f = list(map(float, line.split()))
        X.append(np.array(f[0::2]) + 1j * np.array(f[1::2]))
    return w, np.array(X)


def _matrices():
    rng = np.random.default_rng(7)
    for n in (1, 2, 3, 5, 8, 16, 33, 64):
        yield f"random{n}", rng.normal(size=(n, n))
        d = np.sort(rng.uniform(0.2, 2.0, n))
        yield f"neardiag{n}", np.diag(d) + 0.05 * rng.normal(size=(n, n))
    # a rotation block (complex pair) with real eigenvalues beside it
    yield "rotation", np.array([[0.0, -1.0, 0.3], [1.0, 0.0, 0.2], [0.0, 0.0, 0.5]])
    # a Davidson-like subspace matrix: symmetric part plus a small skew part
    B = rng.normal(size=(24, 24))
    yield "nearsym", np.diag(np.linspace(0.3, 3.0, 24)) + 0.01 * (B + B.T) + 1e-3 * (B - B.T)
    yield "zero", np.zeros((4, 4))


@pytest.mark.parametrize("name,A", list(_matrices()), ids=lambda x: x if isinstance(x, str) else "")
def test_eigenvalues_and_vectors(driver, name, A):
    w, X = _run(driver, A)
    ref = np.linalg.eigvals(A)
    scale = max(1.0, np.abs(ref).max())
    # every computed value matches one of numpy's and vice versa
    assert max(np.abs(ref - x).min() for x in w) / scale < 1e-11
    assert max(np.abs(w - r).min() for r in ref) / scale < 1e-11
    for z, x in zip(w, X):
        assert abs(np.linalg.norm(x) - 1.0) < 1e-12
        assert np.linalg.norm(A @ x - z * x) / scale < 1e-9
