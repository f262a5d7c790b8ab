"""The small dense eigensolvers' arithmetic (eigenfns_amd/csrc/heevql_kernels.hpp)
run serially on the host: the header's per-matrix bodies — Householder + QL
(heev_ql_body) and Householder + bisection + inverse iteration (heevx_body) —
compiled with g++ for a team of one thread (the device kernels run the same
code with 256 threads and barriers between their phases), against LAPACK
(numpy.linalg.eigh) on the same FP64 matrices.

Bars: eigenvalues within 1e-13·max|A| of LAPACK's, residual ‖AV − VΛ‖ within
1e-12·max|A|, ‖VᴴV − I‖ below 1e-12, zero padding past the problem's size;
random, repeated-eigenvalue, close-cluster, tight-cluster, diagonal, zero, penalised
(Rayleigh–Ritz with dead rows) and isolated (dead rows decoupled) matrices of sizes 1-96, every lowest subset
the leading columns of the full solve bit for bit."""
import ctypes as C
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

METHODS = {"ql": 0, "heevx": 1}
HEADER = Path(__file__).resolve().parents[1] / "eigenfns_amd" / "csrc" / "heevql_kernels.hpp"
SOURCE = """
#include "%s"
#include <vector>
extern "C" int heev_host(int method, int n, int lda, const double* A, int lowest, double* w, double* V) {
  std::vector<char> lds(method == 0 ? efn::ql::Layout(lda).bytes() : efn::ql::heevx_lds_bytes(lda));
  efn::ql::Team t{0, 1};
  int info = -1;
  const auto* a = reinterpret_cast<const efn::ql::cd*>(A);
  auto* v = reinterpret_cast<efn::ql::cd*>(V);
  if (method == 0)
    efn::ql::heev_ql_body(t, a, lda, n, lda, lowest, w, v, &info, lds.data());
  else
    efn::ql::heevx_body(t, a, lda, n, lda, lowest, w, v, &info, lds.data());
  return info;
}
"""


@pytest.fixture(scope="module")
def ql(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("ql")
    (d / "ql.cpp").write_text(SOURCE % HEADER)
    subprocess.run([gxx, "-O2", "-std=c++17", "-shared", "-fPIC", "-o", str(d / "ql.so"), str(d / "ql.cpp")],
                   check=True)
    lib = C.CDLL(str(d / "ql.so"))
    lib.heev_host.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]

    def run(method, A, lowest, lda):
        n = A.shape[0]
        Ap = np.zeros((lda, lda), np.complex128)
        Ap[:n, :n] = A
        w = np.empty(lowest)
        V = np.empty((lda, lowest), np.complex128)
        info = lib.heev_host(METHODS[method], n, lda, Ap.ctypes.data, lowest, w.ctypes.data, V.ctypes.data)
        return info, w, V
    return run


def _matrix(rng, n, kind):
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    if kind == "random":
        A = X + X.conj().T
    elif kind == "repeated":  # every eigenvalue three times
        Q, _ = np.linalg.qr(X)
        A = (Q * np.repeat(rng.standard_normal((n + 2) // 3), 3)[:n]) @ Q.conj().T
    elif kind == "close":  # clusters of four eigenvalues 1e-10 apart
        Q, _ = np.linalg.qr(X)
        lam = np.repeat(rng.standard_normal((n + 3) // 4), 4)[:n] + 1e-10 * np.tile(np.arange(4), n)[:n]
        A = (Q * lam) @ Q.conj().T
    elif kind == "tight":  # up to 30 eigenvalues 1e-13 apart, the rest spread
        Q, _ = np.linalg.qr(X)
        lam = rng.standard_normal(n)
        lam[:30] = 1.0 + 1e-13 * np.arange(min(30, n))
        A = (Q * lam) @ Q.conj().T
    elif kind == "diagonal":
        A = np.diag(rng.standard_normal(n)).astype(np.complex128)
    elif kind == "zero":
        A = np.zeros((n, n), np.complex128)
    elif kind == "isolated":  # dead rows decoupled exactly, penalties 1e4·(1 + 1e-6·j) (the solver's form)
        A = X @ X.conj().T / n
        dead = np.flatnonzero(rng.random(n) < 0.3)
        A[dead, :] = 0
        A[:, dead] = 0
        A[dead, dead] = 1e4 * (1.0 + 1e-6 * np.arange(len(dead)))
    else:  # PSD with dead rows penalised at 1e4 (a Rayleigh–Ritz matrix)
        A = X @ X.conj().T / n
        dead = rng.random(n) < 0.3
        A[np.diag_indices(n)] += np.where(dead, 1e4 * (1.0 + 1e-3 * np.arange(n)), 0.0)
    return (A + A.conj().T) / 2


@pytest.mark.parametrize("method", ["ql", "heevx"])
@pytest.mark.parametrize("kind", ["random", "repeated", "close", "tight", "diagonal", "zero", "penalised", "isolated"])
@pytest.mark.parametrize("n", [1, 2, 3, 5, 17, 33, 72, 96])
def test_body_matches_lapack(ql, n, kind, method):
    rng = np.random.default_rng(1000 * n + len(kind))
    A = _matrix(rng, n, kind)
    lda = min(n + 3, 96) if n < 96 else 96
    ref = np.linalg.eigvalsh(A)
    scale = max(np.abs(A).max(), 1e-300)
    info, wf, Vf = ql(method, A, n, lda)
    assert info == 0
    for low in sorted({1, max(1, n // 3), n}):
        info, w, V = ql(method, A, low, lda)
        assert info == 0
        assert np.abs(w - ref[:low]).max() <= 1e-13 * scale
        Vn = V[:n]
        assert np.abs(A @ Vn - Vn * w[None, :]).max() <= 1e-12 * scale
        assert np.abs(Vn.conj().T @ Vn - np.eye(low)).max() < 1e-12
        assert not V[n:].any()
        assert np.array_equal(w, wf[:low]) and np.array_equal(V, Vf[:, :low])
