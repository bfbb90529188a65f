"""CPU restatement of the cosine-transform direct solver.  Test
infrastructure, never part of the product.

Follows init_FFT (code/handwritten/FFT_solver/FFT_solver_extend.f90:43-126),
solve_FFT_SF_x/y/z (code/handwritten/FFT_solver/FFT_algorithm.f90:27-210),
dct1D (code/handwritten/ops/ops_dct.f90:36-70) and idct1D
(code/handwritten/ops/ops_idct.f90:36-71): every plane normal to `direction`
is an independent 2D problem; the interior is transformed by the staggered
DCT-II along the two transverse directions, each mode (p, q) divided by the
eigenvalue of the 5-point Neumann Laplacian, the (0, 0) mode set to zero, and
transformed back.  The eigenvalue is the exact one for any uniform transverse
spacings; the reference's factor 0.5 dy dz / (cos + cos - 2) equals it when
the two spacings are equal, and for direction 1 the reference guards the zero
mode with the indices (i, j) rather than (j, k).

The DCT is restated twice: as the cosine matrices the solver uses, and as the
reference's own construction (mirror extension, FFT of length 2N, twiddle,
real part), so that tests can pin one against the other.
"""
from __future__ import annotations

import numpy as np


def dct_matrix(N: int) -> np.ndarray:
    """C[p, j] = cos(pi p (j + 1/2) / N): Y(p) = sum_j X(j) C[p, j] (ops_dct.f90:40-44)."""
    p = np.arange(N)[:, None]
    j = np.arange(N)[None, :]
    return np.cos(np.pi * p * (j + 0.5) / N)


def idct_matrix(N: int) -> np.ndarray:
    """Ci[j, p] = w_p cos(pi p (j + 1/2) / N), w_0 = 1/N, w_p = 2/N."""
    w = np.full(N, 2.0 / N)
    w[0] = 1.0 / N
    return dct_matrix(N).T * w[None, :]


def dct_reference_construction(x: np.ndarray) -> np.ndarray:
    """dct1D as the reference computes it (ops_dct.f90:46-69): xx = [x, x
    reversed], FFT of length 2N, times e(k) = 0.5 exp(-i pi (k-1) / (2N)),
    real part of the first N entries."""
    N = x.shape[0]
    xx = np.concatenate([x, x[::-1]]).astype(complex)
    X = np.fft.fft(xx)
    e = 0.5 * np.exp(-1j * 0.5 * np.pi * np.arange(N) / N)
    return np.real(X[:N] * e)


def idct_reference_construction(y: np.ndarray) -> np.ndarray:
    """idct1D as the reference computes it (ops_idct.f90:47-70): xx = y e,
    xx2 = [xx, 0, conj(xx(N..2))], FFT of length 2N, 2/N times the real part
    of the first N entries."""
    N = y.shape[0]
    e = 0.5 * np.exp(-1j * 0.5 * np.pi * np.arange(N) / N)
    xx = y.astype(complex) * e
    xx2 = np.zeros(2 * N, dtype=complex)
    xx2[:N] = xx
    for k in range(2, N + 1):  # xx2(N+k) = conjg(xx(N-k+2)), 1-based
        xx2[N + k - 1] = np.conj(xx[N - k + 1])
    X = np.fft.fft(xx2)
    return 2.0 / N * np.real(X[:N])


def transverse(direction: int):
    d = direction - 1
    return d, [t for t in range(3) if t != d]


def inv_eigenvalues(n, h, a, b) -> np.ndarray:
    """1/lambda(p, q) on the (Na, Nb) mode grid, 0 at (0, 0)."""
    Na, Nb = n[a], n[b]
    lp = (2.0 * np.cos(np.pi * np.arange(Na) / Na) - 2.0) / h[a] ** 2
    lq = (2.0 * np.cos(np.pi * np.arange(Nb) / Nb) - 2.0) / h[b] ** 2
    lam = lp[:, None] + lq[None, :]
    out = np.zeros_like(lam)
    nz = lam != 0.0
    out[nz] = 1.0 / lam[nz]
    out[0, 0] = 0.0
    return out


def reference_factors(n, h_first, direction: int) -> np.ndarray:
    """init_FFT's mode factors (code/handwritten/FFT_solver/FFT_solver_extend.f90:62-112)
    on the full interior (n0, n1, n2): 0.5 ha hb / (cos(pi p/Na) + cos(pi q/Nb) - 2)
    with ha, hb the first spacings dhn(1) of the transverse directions; for
    direction 1 the guard tests (i, j), so the modes (0, q) of the first
    x-plane get 1; the (0, 0) modes, set to zero right after ("pin down
    pressure", FFT_algorithm.f90:68-71), are 0 here."""
    d, (a, b) = transverse(direction)
    Na, Nb = n[a], n[b]
    ca = np.cos(np.pi * np.arange(Na) / Na)
    cb = np.cos(np.pi * np.arange(Nb) / Nb)
    with np.errstate(divide="ignore"):
        c2 = 0.5 * h_first[a] * h_first[b] / (ca[:, None] + cb[None, :] - 2.0)
    c2[0, 0] = 0.0
    shape = [1, 1, 1]
    shape[a], shape[b] = Na, Nb
    c = np.broadcast_to(c2.reshape(shape), n).copy()
    if direction == 1:
        c[0, 0, 1:] = 1.0
    return c


def solve_interior(f_int: np.ndarray, h, direction: int, h_first=None) -> np.ndarray:
    """u's interior (n0, n1, n2) from f's interior.  h_first given: the
    reference's mode factors from the first spacings; else the exact
    eigenvalues for the transverse spacings h."""
    n = f_int.shape
    _, (a, b) = transverse(direction)
    y = np.moveaxis(np.tensordot(dct_matrix(n[a]), f_int, axes=([1], [a])), 0, a)
    y = np.moveaxis(np.tensordot(dct_matrix(n[b]), y, axes=([1], [b])), 0, b)
    if h_first is not None:
        y = y * reference_factors(n, h_first, direction)
    else:
        c = inv_eigenvalues(n, h, a, b)
        shape = [1, 1, 1]
        shape[a], shape[b] = n[a], n[b]
        y = y * c.reshape(shape)
    y = np.moveaxis(np.tensordot(idct_matrix(n[a]), y, axes=([1], [a])), 0, a)
    y = np.moveaxis(np.tensordot(idct_matrix(n[b]), y, axes=([1], [b])), 0, b)
    return y
