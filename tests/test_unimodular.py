"""The unimodular completion of the preparation (pfvs_amd/csrc/pfa_prep.h,
the orthogonal-lattice step of pfvs/fp_kernel/pfv_lattice.h:151-188):
compiled for the host with g++ (tests/unimodular_check.cpp), the reduced
Euclid chain that takes over when the plain chain overflows int64 is
checked exactly with Python integers on vectors up to 2^52 -- U v = (g, 0,
..., 0) with |g| = gcd(v) and g equal to the plain chain's, |det U| = 1 (so
rows 1.. are a basis of the lattice orthogonal to v), entries within
n max|v| -- and it is exercised on vectors where the plain chain overflows."""

import math
import os
import shutil
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def det_bareiss(A):
    """Exact determinant (fraction-free Gaussian elimination, Python ints)."""
    A = [list(r) for r in A]
    n, sign, prev = len(A), 1, 1
    for k in range(n - 1):
        if A[k][k] == 0:
            sw = next((i for i in range(k + 1, n) if A[i][k] != 0), None)
            if sw is None:
                return 0
            A[k], A[sw] = A[sw], A[k]
            sign = -sign
        for i in range(k + 1, n):
            for j in range(k + 1, n):
                A[i][j] = (A[i][j] * A[k][k] - A[i][k] * A[k][j]) // prev
        prev = A[k][k]
    return sign * A[n - 1][n - 1]


def test_det_bareiss_pins():
    assert det_bareiss([[2, 1], [7, 4]]) == 1
    assert det_bareiss([[0, 1, 0], [1, 0, 0], [0, 0, 1]]) == -1
    assert det_bareiss([[1, 2], [2, 4]]) == 0
    assert det_bareiss([[3, 0, 0], [5, 2, 0], [1, 1, 7]]) == 42


def _cases():
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    with tempfile.TemporaryDirectory() as td:
        exe = os.path.join(td, "unimodular_check")
        subprocess.run([gxx, "-O2", "-std=c++17", "-ffp-contract=off", "-o", exe,
                        os.path.join(HERE, "unimodular_check.cpp")], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, timeout=120, check=True)
    for line in out.stdout.splitlines():
        n, v, rc0, rc1, U, g0 = line.split("|")
        n = int(n)
        yield (n, [int(x) for x in v.split()], int(rc0), int(rc1),
               [int(x) for x in U.split()], int(g0))


def test_reduced_chain_is_a_unimodular_completion():
    n_cases = n_over = 0
    for n, v, rc0, rc1, U, g0 in _cases():
        assert rc1 == 0, (n, v)
        rows = [U[i * n:(i + 1) * n] for i in range(n)]
        Uv = [sum(a * b for a, b in zip(r, v)) for r in rows]
        g = math.gcd(*v)
        assert abs(Uv[0]) == g and all(x == 0 for x in Uv[1:]), (n, v)
        assert abs(det_bareiss(rows)) == 1, (n, v)
        assert max(abs(x) for x in U) <= n * max(1, max(abs(x) for x in v)), (n, v)
        if rc0 == 0:
            assert Uv[0] == g0, (n, v)          # the same gcd step sequence
        else:
            n_over += 1
        n_cases += 1
    assert n_cases >= 100 and n_over >= 30
