"""The banded LU scheme of csrc/band.hip restated in numpy and checked
against dense solves (CPU): band storage with KV = kl + ku superdiagonals
for the interchanges' fill and BNB extra rows for a panel's multipliers in
the panel's permuted order, partial pivoting inside the panel, the panel's
interchanges and unit-lower triangle on the trailing columns, the masked
rank-BNB update of the kl x KV window, and the solves panel by panel.  It
pins the index arithmetic the kernels share (which entries lie inside the
band, where a panel's rows and columns start and end, the last panel
narrower than BNB)."""
import numpy as np
import pytest
import scipy.sparse as sp

NB = 16  # the kernels' panel width (BNB)

def band_from_csr(A, kl, ku):
    N = A.shape[0]; KV = kl + ku; LD = 2 * kl + ku + NB
    AB = np.zeros((N, LD))  # AB[col, KV + row - col]
    coo = A.tocoo()
    AB[coo.col, KV + coo.row - coo.col] = coo.data
    return AB, KV, LD

def factor(AB, N, kl, KV):
    ipiv = np.zeros(N, np.int64)
    for j0 in range(0, N, NB):
        jb = min(NB, N - j0)
        rmax = min(N, j0 + jb + kl)
        # panel: rows j0..rmax-1, cols j0..j0+jb-1 into a dense array
        P = np.zeros((rmax - j0, jb))
        for k in range(jb):
            c = j0 + k
            for ii in range(rmax - j0):
                i = j0 + ii
                if i - c <= kl + NB - 1 and c - i <= KV:
                    P[ii, k] = AB[c, KV + i - c]
        for k in range(jb):
            c = j0 + k
            lo, hi = k, min(rmax - j0, k + kl + 1)
            col = np.abs(P[lo:hi, k])
            p = lo + int(np.argmax(col))
            ipiv[c] = j0 + p
            if p != k:
                tmp = P[k, :].copy(); P[k, :] = P[p, :]; P[p, :] = tmp
            piv = P[k, k]
            if piv != 0:
                P[k + 1:hi, k] /= piv
                P[k + 1:hi, k + 1:] -= np.outer(P[k + 1:hi, k], P[k, k + 1:])
        for k in range(jb):
            c = j0 + k
            for ii in range(rmax - j0):
                i = j0 + ii
                if i - c <= kl + NB - 1 and c - i <= KV:
                    AB[c, KV + i - c] = P[ii, k]
                else:
                    assert P[ii, k] == 0
        if j0 + jb >= N:
            continue
        qmax = min(N, j0 + jb + KV)
        # swaps + TRSM on trailing columns
        for q in range(j0 + jb, qmax):
            for k in range(jb):
                c = j0 + k; p = ipiv[c]
                if q - c <= KV and p != c:
                    a, b = AB[q, KV + c - q], AB[q, KV + p - q]
                    AB[q, KV + c - q], AB[q, KV + p - q] = b, a
            for k in range(jb):
                c = j0 + k
                if q - c > KV:
                    continue
                u = AB[q, KV + c - q]
                for k2 in range(k + 1, jb):
                    AB[q, KV + j0 + k2 - q] -= P[k2, k] * u
        # GEMM update
        r0, r1 = j0 + jb, min(N, j0 + jb + kl)
        for q in range(j0 + jb, qmax):
            U = np.array([AB[q, KV + j0 + k - q] if q - (j0 + k) <= KV else 0.0 for k in range(jb)])
            rows = np.arange(r0, r1)
            Lm = np.zeros((r1 - r0, jb))
            for k in range(jb):
                c = j0 + k
                ok = rows - c <= kl + NB - 1
                Lm[ok, k] = AB[c, KV + rows[ok] - c]
            AB[q, KV + rows - q] -= Lm @ U
    return ipiv

def solve(AB, ipiv, N, kl, KV, b):
    b = b.copy()
    for j0 in range(0, N, NB):
        jb = min(NB, N - j0)
        for j in range(j0, j0 + jb):
            p = ipiv[j]
            if p != j:
                b[j], b[p] = b[p], b[j]
        rmax = min(N, j0 + jb + kl)
        for j in range(j0, j0 + jb):
            b[j + 1:rmax] -= AB[j, KV + np.arange(j + 1, rmax) - j] * b[j]
    for j in range(N - 1, -1, -1):
        b[j] /= AB[j, KV]
        lo = max(0, j - KV)
        b[lo:j] -= AB[j, KV + np.arange(lo, j) - j] * b[j]
    return b



@pytest.mark.parametrize("N,kl,ku,seed", [(40, 3, 2, 1), (200, 17, 9, 2), (333, 20, 20, 3), (97, 31, 5, 4)])
def test_band_lu_scheme_matches_dense_solve(N, kl, ku, seed):
    rng = np.random.default_rng(seed)
    A = sp.diags([rng.standard_normal(N - abs(k)) for k in range(-kl, ku + 1)], list(range(-kl, ku + 1)),
                 format="csr")
    AB, KV, LD = band_from_csr(A, kl, ku)
    ipiv = factor(AB, N, kl, KV)
    b = rng.standard_normal(N)
    x = solve(AB, ipiv, N, kl, KV, b)
    xd = np.linalg.solve(A.toarray(), b)
    assert np.linalg.norm(x - xd) <= 1e-10 * np.linalg.norm(xd)
    # pivots stay inside the band, never above their column
    assert np.all(ipiv >= np.arange(N)) and np.all(ipiv <= np.arange(N) + kl)


def test_band_lu_scheme_saddle_point_needs_pivoting():
    """A Stokes-like block with zero pressure diagonals in node order: the
    scheme's partial pivoting solves it (no-pivot elimination meets a zero
    pivot at the first pressure row)."""
    rng = np.random.default_rng(5)
    nn = 30
    N = 3 * nn
    A = np.zeros((N, N))
    for m in range(nn):
        for c in (1, 2):
            u = 3 * m + c
            A[u, u] = 4.0 + rng.random()
            for d in (-3, 3):
                if 0 <= u + d < N:
                    A[u, u + d] = -1.0
            A[u, 3 * m] = 1.0          # pressure gradient
            A[3 * m, u] = 1.0          # continuity
    As = sp.csr_matrix(A)
    kl = ku = int(np.max(np.abs(As.tocoo().row - As.tocoo().col)))
    AB, KV, LD = band_from_csr(As, kl, ku)
    ipiv = factor(AB, N, kl, KV)
    b = rng.standard_normal(N)
    x = solve(AB, ipiv, N, kl, KV, b)
    assert np.linalg.norm(A @ x - b) <= 1e-10 * np.linalg.norm(b)
    assert np.any(ipiv != np.arange(N))
