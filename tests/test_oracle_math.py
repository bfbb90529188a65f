"""Pins the oracle's operators and PCG by what the mathematics fixes:
closed forms, discrete identities, symmetry of the volume-weighted operator,
dense assembly by unit vectors and a textbook PCG on the assembled matrix."""
import numpy as np
import pytest

from oracle import oracle as O


def grid(Ns=(5, 4, 6), betas=(None, 1.2, 1.05)):
    hs = []
    for N, b in zip(Ns, betas):
        hs.append(O.uniform_nodes(0.0, 1.0, N) if b is None else O.roberts_both_nodes(0.0, 1.0, N, b))
    return O.Grid(*hs)


def coords(g, dl):
    """Physical coordinates of the points of a field at location dl."""
    out = []
    for d in range(3):
        out.append(g.coord(d, 0) if (dl >> d) & 1 else g.coord(d, 1))
    return np.meshgrid(*out, indexing="ij")


def interior(a):
    return a[1:-1, 1:-1, 1:-1]


def test_uniform_spacing_and_cell_centres():
    g = grid((8, 8, 8), (None, None, None))
    dhn = g.coord(0, 2)
    np.testing.assert_allclose(dhn, 1.0 / 8, rtol=0, atol=1e-15)
    np.testing.assert_allclose(g.coord(0, 8), 0.5, atol=1e-15)  # theta on a uniform grid
    np.testing.assert_array_equal(g.coord(0, 4), -g.coord(0, 5))


def test_roberts_grid_properties():
    hn = O.roberts_both_nodes(0.0, 1.0, 32, 1.05)[1:-1]
    assert hn[0] == 0.0 and abs(hn[-1] - 1.0) < 1e-15
    dh = np.diff(hn)
    assert np.all(dh > 0)
    np.testing.assert_allclose(dh, dh[::-1], rtol=1e-12)  # symmetric clustering (alpha = 1/2)
    assert dh[0] < 0.2 * dh[len(dh) // 2]  # clustered at both walls
    # transformation 2 at xbar = 1/2 maps to 1/2
    assert abs(hn[16] - 0.5) < 1e-14


def test_grad_div_exact_on_linear_and_quadratic():
    g = grid()
    X, Y, Z = coords(g, 0)
    u = np.asfortranarray(3.0 * X - 2.0 * Y + 0.5 * Z + 1.0)
    gx, gy, gz = O.grad(g, u, 0)
    # staggered differences of a linear function are exact at interior nodes
    np.testing.assert_allclose(gx[1:-1, 1:-1, 1:-1], 3.0, rtol=1e-12)
    np.testing.assert_allclose(gy[1:-1, 1:-1, 1:-1], -2.0, rtol=1e-12)
    np.testing.assert_allclose(gz[1:-1, 1:-1, 1:-1], 0.5, rtol=1e-12)
    # div of a linear face field is its constant divergence
    f = [np.asfortranarray(c) for c in (lambda Xs: (2 * Xs[0][0], -Xs[1][1], 4 * Xs[2][2]))(
        [coords(g, 1 << d) for d in range(3)])]
    dv = O.div(g, f[0], f[1], f[2], 0)
    np.testing.assert_allclose(interior(dv), 5.0, rtol=1e-12)
    assert np.all(dv[0] == 0) and np.all(dv[-1] == 0)  # ghosts zeroed (no BCs)


def test_lap_exact_on_quadratic_uniform():
    g = grid((6, 7, 5), (None, None, None))
    X, Y, Z = coords(g, 0)
    u = np.asfortranarray(X ** 2 + 2 * Y ** 2 - Z ** 2 + X * Y)
    L = O.lap_centered(g, u, 0)
    np.testing.assert_allclose(interior(L), 2.0 + 4.0 - 2.0, rtol=1e-10)


def test_curl_grad_and_div_curl_vanish():
    rng = np.random.default_rng(1)
    g = grid()
    phi = np.asfortranarray(rng.standard_normal(g.shape(0)))
    gx, gy, gz = O.grad(g, phi, 0)  # faces
    for d, dl_e in enumerate((6, 5, 3)):
        c = O.curl(g, gx, gy, gz, dl_e, d)
        assert np.max(np.abs(c[1:-1, 1:-1, 1:-1])) < 1e-9 * np.max(np.abs(gx))
    # div(curl E) = 0 for any edge field E (edges -> faces -> cells)
    E = [np.asfortranarray(rng.standard_normal(g.shape(dl))) for dl in (6, 5, 3)]
    B = [O.curl(g, E[0], E[1], E[2], 1 << d, d) for d in range(3)]
    dv = O.div(g, B[0], B[1], B[2], 0)
    assert np.max(np.abs(interior(dv))) < 1e-9 * max(np.max(np.abs(b)) for b in B)


def assemble(g, bcs, dl=0):
    """Dense vol*A of the PCG operator (apply implicit BCs, lap, times vol) over interior cells."""
    s = g.shape(dl)
    vol = O.volume(g, dl)
    idx = [(i, j, k) for k in range(1, s[2] - 1) for j in range(1, s[1] - 1) for i in range(1, s[0] - 1)]
    n = len(idx)
    A = np.zeros((n, n))
    for c, (i, j, k) in enumerate(idx):
        e = np.zeros(s, order="F")
        e[i, j, k] = 1.0
        O.apply_bcs(g, e, dl, bcs, implicit=True)
        Ae = O.lap_centered(g, e, dl) * vol
        A[:, c] = [Ae[q] for q in idx]
    return A, idx


@pytest.mark.parametrize("kind", ["Neumann", "Dirichlet", "mixed"])
def test_volume_weighted_laplacian_is_symmetric(kind):
    g = grid()
    types = {"Neumann": ["Neumann"] * 6, "Dirichlet": ["Dirichlet"] * 6,
             "mixed": ["Neumann", "Dirichlet", "Neumann", "Neumann", "Dirichlet", "Neumann"]}[kind]
    A, _ = assemble(g, O.BCs.make(types))
    np.testing.assert_allclose(A, A.T, rtol=0, atol=1e-10 * np.abs(A).max())
    ev = np.linalg.eigvalsh(-A)
    if kind == "Neumann":
        assert abs(ev[0]) < 1e-8 * ev[-1] and ev[1] > 0  # one constant null vector
    else:
        assert ev[0] > 0


def test_jacobi_diagonal_is_dense_diagonal_away_from_walls():
    g = grid()
    bcs = O.BCs.make(["Neumann"] * 6)
    A, idx = assemble(g, bcs)
    diag = O.lap_diagonal(g, 0) * O.volume(g, 0)
    s = g.shape(0)
    for c, (i, j, k) in enumerate(idx):
        wall = i in (1, s[0] - 2) or j in (1, s[1] - 2) or k in (1, s[2] - 2)
        if not wall:
            assert abs(A[c, c] - diag[i, j, k]) <= 1e-12 * abs(A[c, c])
        else:  # documented quirk: laplacian_diagonal ignores the BC-modified wall rows
            assert abs(A[c, c]) < abs(diag[i, j, k])


def _rhs(g, seed, mean_free_vol=True):
    b = np.asfortranarray(O.splitmix(int(np.prod(g.shape(0))), seed).reshape(g.shape(0), order="F"))
    b[0, :, :] = b[-1, :, :] = 0
    b[:, 0, :] = b[:, -1, :] = 0
    b[:, :, 0] = b[:, :, -1] = 0
    return b


@pytest.mark.parametrize("kind", ["Neumann", "Dirichlet"])
def test_pcg_matches_dense_solve(kind):
    g = grid()
    bcs = O.BCs.make([kind] * 6)
    alln = kind == "Neumann"
    b = _rhs(g, 11)
    x = np.zeros(g.shape(0), order="F")
    it, res0, res, hist = O.pcg_lap(g, 0, bcs, alln, x, b, 500, 1e-14, 1e-300)
    assert res / res0 < 1e-13
    A, idx = assemble(g, bcs)
    vol = O.volume(g, 0)
    bb = np.array([b[q] for q in idx])
    if alln:  # singular: the oracle removes the vol-weighted mean of b first
        vv = np.array([vol[q] for q in idx])
        bb = bb - (bb * vv).sum() / vv.sum()
    rhs = bb * np.array([vol[q] for q in idx])
    xd = np.linalg.lstsq(A, rhs, rcond=None)[0]
    xs = np.array([x[q] for q in idx])
    if alln:
        xd -= xd.mean()
        xs -= xs.mean()
    np.testing.assert_allclose(xs, xd, rtol=0, atol=1e-9 * np.abs(xd).max())
    # ghosts of x carry the Neumann/Dirichlet BCs
    if kind == "Neumann":
        np.testing.assert_array_equal(x[0, 1:-1, 1:-1], x[1, 1:-1, 1:-1])
    else:
        np.testing.assert_array_equal(x[0, 1:-1, 1:-1], -x[1, 1:-1, 1:-1])


def test_pcg_residual_history_matches_textbook_pcg():
    """The oracle's per-iteration residuals equal a textbook Jacobi-PCG on the
    assembled matrix with the reference's preconditioner (1/(vol*diag))."""
    g = grid((6, 5, 4), (1.3, None, 1.1))
    bcs = O.BCs.make(["Neumann"] * 6)
    b = _rhs(g, 5)
    x = np.zeros(g.shape(0), order="F")
    it, res0, res, hist = O.pcg_lap(g, 0, bcs, True, x, b, 25, 1e-300, 1e-300)
    A, idx = assemble(g, bcs)
    vol = O.volume(g, 0)
    vv = np.array([vol[q] for q in idx])
    bb = np.array([b[q] for q in idx])
    bb = bb - (bb * vv).sum() / vv.sum()
    r = bb * vv
    minv = 1.0 / np.array([(O.lap_diagonal(g, 0) * vol)[q] for q in idx])
    xs = np.zeros_like(r)
    z = minv * r
    p = z.copy()
    rho = r @ z
    ref = [np.sqrt(abs(rho))]
    for _ in range(25):
        Ap = A @ p
        a = rho / (p @ Ap)
        xs += a * p
        r -= a * Ap
        ref.append(np.linalg.norm(r))
        z = minv * r
        rn = z @ r
        p = z + (rn / rho) * p
        rho = rn
    assert it == 25
    np.testing.assert_allclose(hist, ref[: len(hist)], rtol=1e-8)
    assert abs(res0 - np.linalg.norm(bb * vv)) < 1e-12 * res0


def test_pcg_exit_on_absolute_tolerance_counts_one_iteration():
    g = grid()
    bcs = O.BCs.make(["Neumann"] * 6)
    b = np.zeros(g.shape(0), order="F")
    x = np.zeros(g.shape(0), order="F")
    it, res0, res, hist = O.pcg_lap(g, 0, bcs, True, x, b, 100, 1e-8, 1e-13)
    assert it == 1 and res == 0.0
