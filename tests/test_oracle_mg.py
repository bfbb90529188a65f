"""Pins the multigrid restatement (oracle/mg_oracle.py): its fine-level
operator is the PCG operator divided by vol, the V-cycle is a convergent
stationary iteration, MG-preconditioned CG reaches the dense solution, and
the zebra line smoother solves its lines exactly and follows its level rule."""
import numpy as np
import pytest

from oracle import mg_oracle as MG
from oracle import oracle as O

BCS = {
    "neumann": [2] * 6,
    "duct": [3, 3, 2, 2, 2, 2],
    "dirichlet_mix": [1, 2, 1, 1, 2, 1],
}


def nodes(N, beta=None):
    return O.uniform_nodes(0.0, 1.0, N) if beta is None else O.roberts_both_nodes(0.0, 1.0, N, beta)


@pytest.mark.parametrize("bc", list(BCS))
def test_fine_operator_is_pcg_operator_over_vol(bc):
    hs = [nodes(6, 1.2), nodes(4), nodes(6, 1.05)]
    g = O.Grid(*hs)
    bct = BCS[bc]
    names = {1: "Dirichlet", 2: "Neumann", 3: "Periodic"}
    bcs = O.BCs.make([names[t] for t in bct])
    L = MG.build(hs, bct, max_levels=1)[0]
    vol = O.volume(g, 0)
    s = g.shape(0)
    rng = np.random.default_rng(0)
    for _ in range(3):
        x = np.asfortranarray(rng.standard_normal(s))
        xr = x.copy(order="F")
        O.apply_bcs(g, xr, 0, bcs, implicit=True)
        ref = O.lap_centered(g, xr, 0)[1:-1, 1:-1, 1:-1]
        mine = L.offdiag(x) + L.diag() * x[1:-1, 1:-1, 1:-1]
        np.testing.assert_allclose(mine, ref, rtol=1e-11, atol=1e-11 * np.abs(ref).max())
    assert np.all(vol[1:-1, 1:-1, 1:-1] > 0)


@pytest.mark.parametrize("bc", list(BCS))
@pytest.mark.parametrize("beta", [None, 1.05])
def test_vcycle_contracts(bc, beta):
    N = 32
    hs = [nodes(N), nodes(N, beta), nodes(N, beta)]
    bct = BCS[bc]
    levels = MG.build(hs, bct)
    L = levels[0]
    rng = np.random.default_rng(1)
    f = np.zeros(L.shape())
    f[1:-1, 1:-1, 1:-1] = rng.standard_normal(tuple(L.n))
    if bct[0] != 1 and bct[2] != 1 and bct[4] != 1 and 1 not in bct:  # singular: make f consistent
        vol = np.einsum("i,j,k->ijk", *[L.dhn[d][1:-1] for d in range(3)])
        f[1:-1, 1:-1, 1:-1] -= (f[1:-1, 1:-1, 1:-1] * vol).sum() / vol.sum()
    x = np.zeros(L.shape())
    r0 = np.linalg.norm(L.residual(x, f))
    rates = []
    prev = r0
    for _ in range(6):
        res = np.zeros(L.shape())
        res[1:-1, 1:-1, 1:-1] = L.residual(x, f)
        x += MG.vcycle(levels, res, bct)
        cur = np.linalg.norm(L.residual(x, f))
        rates.append(cur / prev)
        prev = cur
    # point red-black smoothing on uniform cells contracts fast; on the
    # beta=1.05 wall-clustered cells (aspect ratio ~5-10) it still converges
    assert max(rates[1:]) < (0.25 if beta is None else 0.8), rates


def test_mg_pcg_reaches_dense_solution():
    hs = [nodes(8), nodes(8, 1.1), nodes(8, 1.1)]
    bct = BCS["duct"]
    g = O.Grid(*hs)
    bcs = O.BCs.make(["Periodic", "Periodic"] + ["Neumann"] * 4)
    levels = MG.build(hs, bct, coarse_cells=2)
    assert len(levels) == 3
    L = levels[0]
    vol = O.volume(g, 0)[1:-1, 1:-1, 1:-1]
    rng = np.random.default_rng(4)
    b = rng.standard_normal(tuple(L.n))
    b -= (b * vol).sum() / vol.sum()
    r = b * vol  # vol-weighted residual for x0 = 0
    x = np.zeros(tuple(L.n))

    def A(v):
        pad = np.zeros(L.shape())
        pad[1:-1, 1:-1, 1:-1] = v
        return (L.offdiag(pad) + L.diag() * v) * vol

    def V(rr):
        f = np.zeros(L.shape())
        f[1:-1, 1:-1, 1:-1] = rr / vol
        return MG.vcycle(levels, f, bct)[1:-1, 1:-1, 1:-1]

    z = V(r)
    p = z.copy()
    rho = (z * r).sum()
    r0 = np.linalg.norm(r)
    it = 0
    while np.linalg.norm(r) / r0 > 1e-12 and it < 50:
        q = A(p)
        a = rho / (p * q).sum()
        x += a * p
        r -= a * q
        z = V(r)
        beta = -a * (z * q).sum() / rho
        rho = (z * r).sum()
        p = z + beta * p
        it += 1
    assert it < 20, it
    # residual of the original system
    assert np.linalg.norm(A(x) - b * vol) / r0 < 1e-11


def _line_colour_mask(L, d, c):
    """Cells of the lines of colour c along d: (i + k) % 2 == c for y-lines,
    (i + j) % 2 == c for z-lines (1-based indices)."""
    idx = np.meshgrid(*[np.arange(1, k + 1) for k in L.n], indexing="ij")
    others = [e for e in range(3) if e != d]
    return ((idx[others[0]] + idx[others[1]]) % 2) == c


@pytest.mark.parametrize("bc", ["duct", "dirichlet_mix"])
@pytest.mark.parametrize("d", [1, 2])
@pytest.mark.parametrize("c", [0, 1])
def test_line_smooth_solves_its_lines_exactly(bc, d, c):
    """After a zebra line sweep of colour c along d, the residual vanishes on
    every cell of those lines (each line's tridiagonal system is solved with
    the other colour's values on the right-hand side) and the other lines are
    untouched; a zero-guess sweep equals a sweep from x = 0."""
    hs = [nodes(6), nodes(7, 1.2), nodes(5, 1.1)]
    bct = BCS[bc]
    L = MG.build(hs, bct, max_levels=1)[0]
    rng = np.random.default_rng(10 * d + c)
    f = np.zeros(L.shape())
    f[1:-1, 1:-1, 1:-1] = rng.standard_normal(tuple(L.n))
    x = np.zeros(L.shape())
    x[1:-1, 1:-1, 1:-1] = rng.standard_normal(tuple(L.n))
    x0 = x.copy()
    L.line_smooth(x, f, d, c, False)
    mask = _line_colour_mask(L, d, c)
    res = L.residual(x, f)
    assert np.abs(res[mask]).max() <= 1e-12 * np.abs(f).max() * np.abs(L.diag()).max()
    assert np.abs(res[~mask]).max() > 1e-6  # the other colour is not solved
    np.testing.assert_array_equal(x[1:-1, 1:-1, 1:-1][~mask], x0[1:-1, 1:-1, 1:-1][~mask])
    a, b = np.zeros(L.shape()), np.zeros(L.shape())
    L.line_smooth(a, f, d, c, True)
    L.line_smooth(b, f, d, c, False)
    np.testing.assert_array_equal(a, b)


def test_line_directions_follow_the_stretching_rule():
    """Lines along y / z where the cell widths vary by more than 1.5x, the
    direction is not periodic and the level has more than 32^3 cells."""
    hs = [nodes(64), nodes(64, 1.05), nodes(64, 1.05)]
    lv = MG.build(hs, BCS["duct"])
    assert [L.line_dirs for L in lv[:3]] == [[1, 2], [], []]  # 64^3 lines; 32^3 is not above 32^3
    lv = MG.build(hs, [2, 2, 2, 2, 3, 3])  # z periodic
    assert lv[0].line_dirs == [1]
    lv = MG.build([nodes(64)] * 3, BCS["duct"])  # uniform: points
    assert lv[0].line_dirs == []


def _vcycle_rate(levels, bct):
    L = levels[0]
    rng = np.random.default_rng(2)
    f = np.zeros(L.shape())
    f[1:-1, 1:-1, 1:-1] = rng.standard_normal(tuple(L.n))
    vol = np.einsum("i,j,k->ijk", *[L.dhn[d][1:-1] for d in range(3)])
    f[1:-1, 1:-1, 1:-1] -= (f[1:-1, 1:-1, 1:-1] * vol).sum() / vol.sum()
    x = np.zeros(L.shape())
    prev = np.linalg.norm(L.residual(x, f))
    rates = []
    for _ in range(4):
        res = np.zeros(L.shape())
        res[1:-1, 1:-1, 1:-1] = L.residual(x, f)
        x += MG.vcycle(levels, res, bct)
        cur = np.linalg.norm(L.residual(x, f))
        rates.append(cur / prev)
        prev = cur
    return max(rates[1:])


def test_vcycle_with_lines_contracts_faster_on_stretched_cells():
    """On a 48^3 duct clustered at the y and z walls (Roberts beta = 1.05),
    the finest level's zebra lines roughly halve the V-cycle's contraction
    factor against red-black points on the same hierarchy."""
    hs = [nodes(48), nodes(48, 1.05), nodes(48, 1.05)]
    bct = BCS["duct"]
    lines = MG.build(hs, bct)
    assert lines[0].line_dirs == [1, 2]
    points = MG.build(hs, bct)
    points[0].line_dirs = []
    rl, rp = _vcycle_rate(lines, bct), _vcycle_rate(points, bct)
    assert rl < 0.5 and rp > 0.6, (rl, rp)
