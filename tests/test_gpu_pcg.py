"""GPU parity of the FP64 Jacobi-PCG pressure-Poisson solve and clean_div
against the oracle (same seeded inputs), plus size-independent properties at
128^3.  Reductions are summed in a different order on the GPU, so vectors are
compared within 1e-10 relative L2 (the bar for floating-point fields) and
iteration counts exactly."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def nodes(spec):
    return O.uniform_nodes(0.0, 1.0, spec[1]) if spec[0] == "u" else O.roberts_both_nodes(0.0, 1.0, spec[1], spec[2])


CASES = {
    "cube16_neumann": ([("u", 16)] * 3, ["Neumann"] * 6, [0.0] * 6, True),
    "roberts_12x10x14_neumann": ([("r", 12, 1.1), ("u", 10), ("r", 14, 1.05)], ["Neumann"] * 6, [0.0] * 6, True),
    "duct_periodic_x": ([("u", 16), ("r", 12, 1.05), ("r", 12, 1.05)],
                        ["Periodic", "Periodic", "Neumann", "Neumann", "Neumann", "Neumann"], [0.0] * 6, True),
    "dirichlet_inhomog": ([("u", 10), ("r", 9, 1.2), ("u", 11)], ["Dirichlet", "Neumann", "Dirichlet", "Dirichlet",
                                                                    "Neumann", "Dirichlet"],
                          [0.5, 0.1, -0.3, 1.0, 0.0, 0.2], False),
    # thin conducting walls on y and z (c_w 0.01 and 0.5), x periodic
    "robin_thin_wall": ([("u", 12), ("r", 10, 1.05), ("r", 11, 1.1)],
                        ["Periodic", "Periodic", "Robin", "Robin", "Robin", "Neumann"], [0.0, 0.0, 0.1, 0.0, -0.2, 0.0],
                        False, [0.0, 0.0, 0.01, 0.01, 0.5, 0.0]),
}


def rhs(shape, seed):
    b = O.splitmix(int(np.prod(shape)), seed).reshape(shape, order="F")
    return np.asfortranarray(b)


class _Types(list):
    """BC types that carry the faces' c_w into set_BCs."""


def setup(gpu, case):
    specs, types, vals, alln = CASES[case][:4]
    cw = CASES[case][4] if len(CASES[case]) > 4 else None
    hs = [nodes(s) for s in specs]
    g = O.Grid(*hs)
    m = gpu.mesh(*hs)
    if cw is not None:
        types = _Types(types)
        types.c_w = cw
    return g, m, O.BCs.make(types, vals, cw), types, vals, alln


def set_bcs(X, types, vals, alln):
    """set_BCs with the case's c_w (Robin faces)."""
    return X.set_BCs(types, vals, all_Neumann=alln, c_w=getattr(types, "c_w", None))


def rel_l2(a, b):
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("iters", [1, 7, 40])
def test_pcg_fixed_iterations_match_oracle(gpu, case, iters):
    g, m, bcs, types, vals, alln = setup(gpu, case)
    b = rhs(g.shape(0), 3)
    x0 = rhs(g.shape(0), 4) * 0.1
    xr = x0.copy(order="F")
    it_r, res0_r, res_r, _ = O.pcg_lap(g, 0, bcs, alln, xr, b, iters, 0.0, 0.0)
    X = set_bcs(gpu.SF(m, 0).upload(x0), types, vals, alln)
    B = gpu.SF(m, 0).upload(b)
    P = gpu.PCG_solver_SF(m, X, gpu.iter_solver_params(iters, 0.0, 0.0, 1))
    info = P.solve(X, B)
    assert info.iterations == it_r == iters
    assert abs(info.res_norm0 - res0_r) <= 1e-12 * res0_r
    assert abs(info.res_norm - res_r) <= 1e-9 * res_r
    assert rel_l2(X.download(), xr) <= 1e-10


@pytest.mark.parametrize("case", list(CASES))
def test_pcg_converged_matches_oracle(gpu, case):
    g, m, bcs, types, vals, alln = setup(gpu, case)
    b = rhs(g.shape(0), 9)
    xr = np.zeros(g.shape(0), order="F")
    it_r, res0_r, res_r, hist = O.pcg_lap(g, 0, bcs, alln, xr, b, 2000, 1e-8, 1e-300)
    X = set_bcs(gpu.SF(m, 0), types, vals, alln)
    B = gpu.SF(m, 0).upload(b)
    P = gpu.PCG_solver_SF(m, X, gpu.iter_solver_params(2000, 1e-8, 1e-300, 1))
    info = P.solve(X, B)
    assert info.iterations == it_r
    assert info.exit_loop & 1
    assert info.res_norm / info.res_norm0 < 1e-8
    assert rel_l2(X.download(), xr) <= 1e-10


def test_pcg_n_skip_and_iter_max_exits(gpu):
    g, m, bcs, types, vals, alln = setup(gpu, "cube16_neumann")
    b = rhs(g.shape(0), 5)
    for n_skip, it_max, tol in ((5, 1000, 1e-6), (1, 13, 1e-12), (4, 30, 1e-12)):
        xr = np.zeros(g.shape(0), order="F")
        it_r, _, res_r, _ = O.pcg_lap(g, 0, bcs, alln, xr, b, it_max, tol, 1e-300, n_skip)
        X = set_bcs(gpu.SF(m, 0), types, vals, alln)
        P = gpu.PCG_solver_SF(m, X, gpu.iter_solver_params(it_max, tol, 1e-300, n_skip))
        info = P.solve(X, gpu.SF(m, 0).upload(b))
        assert info.iterations == it_r
        assert rel_l2(X.download(), xr) <= 1e-10


def test_pcg_zero_rhs_exits_on_tol_abs(gpu):
    g, m, bcs, types, vals, alln = setup(gpu, "cube16_neumann")
    X = gpu.SF(m, 0).set_BCs(types, vals, all_Neumann=True)
    P = gpu.PCG_solver_SF(m, X, gpu.iter_solver_params(100, 1e-8, 1e-13, 1))
    info = P.solve(X, gpu.SF(m, 0))
    assert info.iterations == 1 and info.exit_loop & 2


def test_clean_div_matches_oracle(gpu):
    """clean_div on a face field: PPE solve + projection, vs the oracle sequence."""
    g, m, bcs, types, vals, alln = setup(gpu, "roberts_12x10x14_neumann")
    rng = np.random.default_rng(2)
    us = [np.asfortranarray(rng.standard_normal(g.shape(1 << d)) * 1e-3) for d in range(3)]
    ubc = (["Dirichlet"] * 6, [0.0, 0.0, 0.0, 1.0, 0.0, 0.0])
    scale = 1.0 / 1e-3
    # oracle sequence (clean_divergence.f90:37-59)
    dv = O.div(g, us[0], us[1], us[2], 0) * scale
    dv = np.asfortranarray(dv)
    phi_r = np.zeros(g.shape(0), order="F")
    O.pcg_lap(g, 0, bcs, True, phi_r, dv, 3000, 1e-12, 1e-300)
    gr = O.grad(g, phi_r, 0)
    Xr = []
    for d in range(3):
        x = np.asfortranarray(us[d] - gr[d] * (1.0 / scale))
        O.apply_bcs(g, x, 1 << d, O.BCs.make(*ubc))
        Xr.append(x)
    # GPU
    Xs = [gpu.SF(m, 1 << d).upload(us[d]) for d in range(3)]
    X = [gpu.SF(m, 1 << d).set_BCs(*ubc) for d in range(3)]
    T = [gpu.SF(m, 1 << d) for d in range(3)]
    phi = gpu.SF(m, 0).set_BCs(types, vals, all_Neumann=True)
    tcc = gpu.SF(m, 0)
    P = gpu.PCG_solver_SF(m, phi, gpu.iter_solver_params(3000, 1e-12, 1e-300, 1))
    info = gpu.clean_div(P, X, Xs, phi, scale, T, tcc)
    assert info.res_norm / info.res_norm0 < 1e-12
    for d in range(3):
        assert rel_l2(X[d].download(), Xr[d]) <= 1e-10
    assert rel_l2(phi.download(), phi_r) <= 1e-10


@pytest.mark.parametrize("N,beta", [(128, None), (128, 1.05)])
def test_pcg_128_true_residual(gpu, N, beta):
    """Full size: the oracle recomputes the true vol-weighted residual of the
    GPU solution; it must equal the recursive residual the GPU reports."""
    hx = O.uniform_nodes(0.0, 1.0, N)
    hy = hx if beta is None else O.roberts_both_nodes(0.0, 1.0, N, beta)
    g = O.Grid(hx, hy, hy)
    m = gpu.mesh(hx, hy, hy)
    types = ["Neumann"] * 6
    b = rhs(g.shape(0), 17)
    X = gpu.SF(m, 0).set_BCs(types, None, all_Neumann=True)
    P = gpu.PCG_solver_SF(m, X, gpu.iter_solver_params(5000, 1e-8, 1e-300, 1))
    info = P.solve(X, gpu.SF(m, 0).upload(b))
    assert info.res_norm / info.res_norm0 < 1e-8
    x = X.download()
    xi = x.copy(order="F")
    O.apply_bcs(g, xi, 0, O.BCs.make(types), True)
    vol = O.volume(g, 0)
    bi = b[1:-1, 1:-1, 1:-1]
    vi = vol[1:-1, 1:-1, 1:-1]
    mean = (bi * vi).sum() / vi.sum()
    r = (bi - mean - O.lap_centered(g, xi, 0)[1:-1, 1:-1, 1:-1]) * vi
    true_rel = np.linalg.norm(r) / info.res_norm0
    assert true_rel < 2e-8, true_rel
    # Neumann ghosts
    np.testing.assert_array_equal(x[0, 1:-1, 1:-1], x[1, 1:-1, 1:-1])
