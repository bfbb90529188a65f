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


def first_spacings(hs):
    return [float(hn[1] - hn[0]) for hn in hs]


@pytest.mark.parametrize("direction", [1, 2, 3])
def test_reference_factors_solve_the_geometric_mean_laplacian(direction):
    """With unequal transverse spacings ha, hb the reference's factor
    0.5 ha hb / (cos + cos - 2) is the inverse eigenvalue of the 5-point
    Laplacian whose two spacings are both sqrt(ha hb): its solution satisfies
    that Laplacian on every plane (the first x-plane aside for direction 1)
    and differs from the exact solution."""
    hs = [O.uniform_nodes(0.0, 1.0, 12), O.uniform_nodes(0.0, 0.7, 10), O.uniform_nodes(0.0, 1.3, 8)]
    g = O.Grid(*hs)
    h = [float(np.diff(hn)[1]) for hn in hs]
    h0 = first_spacings(hs)
    d = direction - 1
    a, b = [t for t in range(3) if t != d]
    rng = np.random.default_rng(10 + direction)
    fi = rng.standard_normal(tuple(s - 2 for s in g.shape(0)))
    ui = FO.solve_interior(fi, h, direction, h_first=h0)
    gm = np.sqrt(h0[a] * h0[b])
    n = [s - 2 for s in g.shape(0)]
    hs2 = list(hs)
    hs2[a] = O.uniform_nodes(0.0, gm * n[a], n[a])
    hs2[b] = O.uniform_nodes(0.0, gm * n[b], n[b])
    g2 = O.Grid(*hs2)
    u = np.zeros(g2.shape(0), order="F")
    u[1:-1, 1:-1, 1:-1] = ui
    O.apply_bcs(g2, u, 0, O.BCs.make(["Neumann"] * 6))
    L = lap_transverse(g2, u, d)[1:-1, 1:-1, 1:-1]
    expect = fi - fi.mean(axis=(a, b), keepdims=True)
    planes = slice(1, None) if direction == 1 else slice(None)
    idx = [slice(None)] * 3
    idx[d] = planes
    idx = tuple(idx)
    assert np.abs(L[idx] - expect[idx]).max() < 1e-10 * np.abs(fi).max()
    exact = FO.solve_interior(fi, h, direction)
    assert np.linalg.norm(exact - ui) > 1e-3 * np.linalg.norm(exact)


def test_reference_factors_equal_exact_for_equal_spacings_but_the_x_quirk():
    """Equal transverse spacings: the reference's factors are the exact
    inverse eigenvalues everywhere, except (direction 1) the modes (0, q > 0)
    of the first x-plane, which the reference's (i, j) guard leaves at 1 --
    there the transform of u equals the transform of f."""
    hs = [O.uniform_nodes(0.0, 1.0, 10), O.uniform_nodes(0.0, 1.0, 12), O.uniform_nodes(0.0, 1.0, 12)]
    h = [float(np.diff(hn)[1]) for hn in hs]
    h0 = first_spacings(hs)
    n = (10, 12, 12)
    for direction in (1, 2, 3):
        d, (a, b) = FO.transverse(direction)
        if abs(h[a] - h[b]) > 0:
            continue
        ref = FO.reference_factors(n, h0, direction)
        ex = np.broadcast_to(FO.inv_eigenvalues(n, h, a, b).reshape([n[t] if t in (a, b) else 1 for t in range(3)]), n)
        diff = np.abs(ref - ex) > 1e-12 * np.abs(ex).max()
        if direction == 1:
            assert diff[0, 0, 1:].all() and diff.sum() == n[2] - 1
            np.testing.assert_array_equal(ref[0, 0, 1:], 1.0)
        else:
            assert not diff.any()
    rng = np.random.default_rng(4)
    fi = rng.standard_normal(n)
    ui = FO.solve_interior(fi, h, 1, h_first=h0)
    C1, C2 = FO.dct_matrix(n[1]), FO.dct_matrix(n[2])
    fu = C1 @ ui[0] @ C2.T
    ff = C1 @ fi[0] @ C2.T
    np.testing.assert_allclose(fu[0, 1:], ff[0, 1:], rtol=1e-12, atol=1e-12 * np.abs(ff).max())
