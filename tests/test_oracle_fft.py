"""The cosine-transform solver's restatement, pinned without the GPU:

* its cosine matrices equal the reference's own DCT construction (mirror
  extension, FFT of length 2N, twiddle, real part; ops_dct.f90:46-69 and
  ops_idct.f90:47-70, restated with numpy.fft) and the textbook DCT-II
  (scipy.fft.dct type 2 is twice the reference's transform);
* the inverse undoes the forward transform;
* the solution satisfies the 5-point Neumann Laplacian in the two transverse
  directions, built from the C oracle's staggered derivatives, with right-hand
  side f minus its mean over each plane, for unequal transverse spacings and
  each direction;
* for equal transverse spacings its mode factors equal the reference's
  0.5 dy dz / (cos + cos - 2) (FFT_solver_extend.f90:67-120)."""
import numpy as np
import pytest
import scipy.fft

from oracle import fft_oracle as FO
from oracle import oracle as O


@pytest.mark.parametrize("N", [4, 8, 12, 16, 32])
def test_dct_matrix_equals_reference_construction(N):
    rng = np.random.default_rng(N)
    for _ in range(3):
        x = rng.standard_normal(N)
        y = FO.dct_matrix(N) @ x
        np.testing.assert_allclose(y, FO.dct_reference_construction(x), rtol=0, atol=1e-12 * np.abs(x).sum())
        np.testing.assert_allclose(y, 0.5 * scipy.fft.dct(x, type=2), rtol=0, atol=1e-12 * np.abs(x).sum())
        np.testing.assert_allclose(FO.idct_matrix(N) @ y, FO.idct_reference_construction(y), rtol=0,
                                   atol=1e-12 * np.abs(y).sum())
        np.testing.assert_allclose(FO.idct_matrix(N) @ y, x, rtol=0, atol=1e-12 * np.abs(x).sum())


def lap_transverse(g, u, d):
    """sum over the two directions transverse to d of del(del(u, CC->N), N->CC)."""
    out = g.zeros(0)
    first = True
    for e in range(3):
        if e == d:
            continue
        t = g.zeros(1 << e)
        O.del_op(g, t, 1 << e, u, 0, 1, e, 0, 1)
        O.del_op(g, out, 0, t, 1 << e, 1, e, 0, 1 if first else 2)
        first = False
    return out


@pytest.mark.parametrize("direction", [1, 2, 3])
def test_solution_satisfies_the_plane_laplacian(direction):
    hs = [O.uniform_nodes(0.0, 1.0, 12), O.uniform_nodes(0.0, 0.7, 10), O.uniform_nodes(0.0, 1.3, 8)]
    g = O.Grid(*hs)
    h = [float(np.diff(hn)[1]) for hn in hs]
    rng = np.random.default_rng(direction)
    f = np.zeros(g.shape(0), order="F")
    f[1:-1, 1:-1, 1:-1] = rng.standard_normal(tuple(s - 2 for s in g.shape(0)))
    u = np.zeros(g.shape(0), order="F")
    u[1:-1, 1:-1, 1:-1] = FO.solve_interior(f[1:-1, 1:-1, 1:-1], h, direction)
    O.apply_bcs(g, u, 0, O.BCs.make(["Neumann"] * 6))
    d = direction - 1
    L = lap_transverse(g, u, d)[1:-1, 1:-1, 1:-1]
    fi = f[1:-1, 1:-1, 1:-1]
    axes = tuple(t for t in range(3) if t != d)
    expect = fi - fi.mean(axis=axes, keepdims=True)
    assert np.abs(L - expect).max() < 1e-10 * np.abs(fi).max()
    # zero mean over every plane (the pinned (0, 0) mode)
    assert np.abs(u[1:-1, 1:-1, 1:-1].mean(axis=axes)).max() < 1e-12 * np.abs(u).max()


def test_equal_spacing_factor_equals_reference_formula():
    n = [16, 16, 8]
    h = [0.1, 0.1, 0.2]
    c = FO.inv_eigenvalues(n, h, 0, 1)
    ci = np.cos(np.pi * np.arange(16) / 16)
    den = ci[:, None] + ci[None, :] - 2.0
    den[0, 0] = 1.0  # the (0, 0) mode is pinned to zero, not divided
    ref = 0.5 * h[0] * h[1] / den
    ref[0, 0] = 0.0
    np.testing.assert_allclose(c, ref, rtol=1e-13, atol=0)
