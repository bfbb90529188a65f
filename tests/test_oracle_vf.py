"""Pins the oracle's vector-field PCG (momentum and induction implicit
diffusion): the volume-weighted operator with the reference's wall factors is
symmetric on the unknowns, the Jacobi diagonal matches the dense diagonal away
from walls, and the PCG reaches the dense solution."""
import numpy as np
import pytest

from oracle import oracle as O

LDC_U = [  # lid-driven cavity velocity: no-slip everywhere, u_x = 1 on the lid (ymax)
    O.BCs.make(["Dirichlet"] * 6, [0, 0, 0, 1.0, 0, 0]),
    O.BCs.make(["Dirichlet"] * 6),
    O.BCs.make(["Dirichlet"] * 6),
]
PSEUDO_VACUUM = [  # B: Neumann for the normal component on its own faces, else Dirichlet 0
    O.BCs.make(["Neumann", "Neumann"] + ["Dirichlet"] * 4),
    O.BCs.make(["Dirichlet"] * 2 + ["Neumann", "Neumann"] + ["Dirichlet"] * 2),
    O.BCs.make(["Dirichlet"] * 4 + ["Neumann", "Neumann"]),
]


def grid(N=(4, 5, 3)):
    return O.Grid(O.uniform_nodes(0, 1, N[0]), O.roberts_both_nodes(0, 1, N[1], 1.2), O.uniform_nodes(0, 1, N[2]))


def unknowns(g, bcs3):
    """Interior points (pad 1) minus Dirichlet wall-node planes, per component."""
    idx = []
    for d in range(3):
        s = g.shape(1 << d)
        for k in range(1, s[2] - 1):
            for j in range(1, s[1] - 1):
                for i in range(1, s[0] - 1):
                    ii = (i, j, k)
                    wall = False
                    for f in range(6):
                        if f // 2 != d:
                            continue
                        plane = 1 if f % 2 == 0 else s[d] - 2
                        if ii[d] == plane and bcs3[d].type[f] == 1:
                            wall = True
                    if not wall:
                        idx.append((d, ii))
    return idx


def assemble(g, kind, c, bcs3, k=None):
    idx = unknowns(g, bcs3)
    vol = [O.volume(g, 1 << d) for d in range(3)]
    A = np.zeros((len(idx), len(idx)))
    for col, (d, ii) in enumerate(idx):
        x = [g.zeros(1 << e) for e in range(3)]
        x[d][ii] = 1.0
        Ax = O.op_vf(g, x, kind, c, bcs3, k)
        for e in range(3):
            s = g.shape(1 << e)
            for f in range(6):
                if f // 2 != e:
                    continue
                plane = 1 if f % 2 == 0 else s[e] - 2
                sl = [slice(None)] * 3
                sl[e] = plane
                if bcs3[e].type[f] == 2:
                    Ax[e][tuple(sl)] *= 0.5
                elif bcs3[e].type[f] == 1:
                    Ax[e][tuple(sl)] = 0.0
            Ax[e] = Ax[e] * vol[e]
        A[:, col] = [Ax[e][jj] for (e, jj) in idx]
    return A, idx


def sigma_inv(g, seed=0):
    rng = np.random.default_rng(seed)
    return [np.asfortranarray(1.0 + 0.5 * rng.random(g.shape(dl))) for dl in (6, 5, 3)]


@pytest.mark.parametrize("kind,bcname", [(1, "ldc"), (2, "pv"), (1, "pv")])
def test_vf_operator_symmetric(kind, bcname):
    g = grid()
    bcs3 = LDC_U if bcname == "ldc" else PSEUDO_VACUUM
    hom = [O.BCs.make([t for t in b.type]) for b in bcs3]
    k = sigma_inv(g) if kind == 2 else None
    # momentum: c = -theta*dt/Re (A = I - theta dt/Re lap); induction: c = +dt/Rem
    c = -0.03 if kind == 1 else 0.03
    A, _ = assemble(g, kind, c, hom, k)
    np.testing.assert_allclose(A, A.T, rtol=0, atol=1e-11 * np.abs(A).max())
    assert np.linalg.eigvalsh(A).min() > 0  # SPD after vol weighting


@pytest.mark.parametrize("kind", [1, 2])
def test_vf_pcg_matches_dense(kind):
    g = grid()
    bcs3 = LDC_U if kind == 1 else PSEUDO_VACUUM
    hom = [O.BCs.make([t for t in b.type]) for b in bcs3]
    k = sigma_inv(g) if kind == 2 else None
    c = -0.05 if kind == 1 else 0.05
    rng = np.random.default_rng(3)
    b = [np.asfortranarray(rng.standard_normal(g.shape(1 << d))) for d in range(3)]
    x = [g.zeros(1 << d) for d in range(3)]
    it, res0, res, _ = O.pcg_vf(g, kind, c, hom, x, b, 500, 1e-13, 1e-300, k=k)
    assert res / res0 < 1e-13
    A, idx = assemble(g, kind, c, hom, k)
    vol = [O.volume(g, 1 << d) for d in range(3)]
    rhs = np.array([b[e][jj] * vol[e][jj] for (e, jj) in idx])
    # wall factors on the right-hand side as in modify_RHS_VF
    for q, (e, jj) in enumerate(idx):
        for f in range(6):
            if f // 2 == e and hom[e].type[f] == 2 and jj[e] == (1 if f % 2 == 0 else g.shape(1 << e)[e] - 2):
                rhs[q] *= 0.5
    xd = np.linalg.solve(A, rhs)
    xs = np.array([x[e][jj] for (e, jj) in idx])
    np.testing.assert_allclose(xs, xd, rtol=0, atol=1e-9 * np.abs(xd).max())


def test_inhomogeneous_lid_enters_through_rhs():
    """With the lid BC u_x = 1 and b = vol*x0-free data, the solution carries the lid value."""
    g = grid((6, 6, 4))
    c = -0.02
    b = [g.zeros(1 << d) for d in range(3)]
    x = [g.zeros(1 << d) for d in range(3)]
    O.pcg_vf(g, 1, c, LDC_U, x, b, 500, 1e-12, 1e-300)
    s = g.shape(1)
    # ghost of u_x above the lid: 2*1 - interior
    np.testing.assert_allclose(x[0][1:-1, -1, 1:-1], 2.0 - x[0][1:-1, -2, 1:-1], atol=1e-14)
    assert x[0][2:-2, -2, 1:-1].min() > 0.0  # lid drags the fluid below it
