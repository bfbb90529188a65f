"""Pins the multigrid restatement (oracle/mg_oracle.py): its fine-level
operator is the PCG operator divided by vol, the V-cycle is a convergent
stationary iteration, and MG-preconditioned CG reaches the dense solution."""
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
