"""Pins the oracle's time-step pieces with what the mathematics fixes:
u.grad(u) and the collocated Laplacian are exact on linear / quadratic fields
in the interior of a uniform grid, J x B has the right sign and pairing, the CT
source is divergence-free, the projection leaves a divergence-free velocity,
and the state at rest with B = 0 stays at rest."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import step_oracle as S

LDC_U = [O.BCs.make(["Dirichlet"] * 6, [0, 0, 0, 1.0, 0, 0]), O.BCs.make(["Dirichlet"] * 6),
         O.BCs.make(["Dirichlet"] * 6)]
P_BCS = O.BCs.make(["Neumann"] * 6)
PV_B = [O.BCs.make(["Neumann", "Neumann"] + ["Dirichlet"] * 4),
        O.BCs.make(["Dirichlet"] * 2 + ["Neumann", "Neumann"] + ["Dirichlet"] * 2),
        O.BCs.make(["Dirichlet"] * 4 + ["Neumann", "Neumann"])]
PHI_BCS = O.BCs.make(["Dirichlet"] * 6)


def coords(g, dl):
    """Coordinates of the points of a field of location dl (hn along N directions, hc otherwise)."""
    out = []
    for d in range(3):
        out.append(g.coord(d, 0) if (dl >> d) & 1 else g.coord(d, 1))
    return np.meshgrid(*out, indexing="ij")


def uniform_grid(N=8):
    return O.Grid(*[O.uniform_nodes(0.0, 1.0, N)] * 3)


def test_convection_exact_on_linear_flow():
    g = uniform_grid()
    U = []
    for d, f in enumerate((lambda x, y, z: x, lambda x, y, z: -y, lambda x, y, z: 0 * z)):
        X, Y, Z = coords(g, 1 << d)
        U.append(np.asfortranarray(f(X, Y, Z)))
    UE = S.face2edge_no_diag(g, U)
    div = S.advect_U_convection(g, U, UE)
    X, Y, Z = coords(g, 1)
    np.testing.assert_allclose(div[0][2:-2, 2:-2, 2:-2], X[2:-2, 2:-2, 2:-2], atol=1e-13)
    X, Y, Z = coords(g, 2)
    np.testing.assert_allclose(div[1][2:-2, 2:-2, 2:-2], Y[2:-2, 2:-2, 2:-2], atol=1e-13)
    np.testing.assert_allclose(div[2][2:-2, 2:-2, 2:-2], 0.0, atol=1e-13)


def test_collocated_laplacian_exact_on_quadratics():
    g = O.Grid(O.roberts_both_nodes(0, 1, 10, 1.2), O.uniform_nodes(0, 1, 9), O.roberts_both_nodes(0, 2, 8, 1.1))
    U = []
    for d in range(3):
        X, Y, Z = coords(g, 1 << d)
        U.append(np.asfortranarray(X * X + 2 * Y * Y - 0.5 * Z * Z + X * Y))
    lap = S.lap_collocated(g, U)
    for d in range(3):
        np.testing.assert_allclose(lap[d][2:-2, 2:-2, 2:-2], 2 + 4 - 1.0, rtol=1e-10)


def test_jcrossb_sign_and_pairing():
    """B = (0, 0, y): J = curl(B) = (1, 0, 0); J x B0 with B0 = (0, 1, 0) is (0, 0, 1)."""
    g = uniform_grid()
    st = S.MHDState.__new__(S.MHDState)
    st.g, st.gi, st.Rem, st.N = g, g, 1.0, 1.0
    st.B0 = [np.asfortranarray(np.full(g.shape(1 << d), (0.0, 1.0, 0.0)[d])) for d in range(3)]
    X, Y, Z = coords(g, 4)
    st.B = [g.zeros(1), g.zeros(2), np.asfortranarray(Y.copy())]
    st.post_induction()
    np.testing.assert_allclose(st.J[0][2:-2, 2:-2, 2:-2], 1.0, atol=1e-12)
    np.testing.assert_allclose(st.jCrossB[2][2:-2, 2:-2, 2:-2], 1.0, atol=1e-12)
    np.testing.assert_allclose(st.jCrossB[0][2:-2, 2:-2, 2:-2], 0.0, atol=1e-12)
    np.testing.assert_allclose(st.jCrossB[1][2:-2, 2:-2, 2:-2], 0.0, atol=1e-12)


def noise_U(g, seed, amp=1e-3):
    U = []
    for d in range(3):
        u = np.asfortranarray(amp * (2 * np.random.default_rng(seed + d).random(g.shape(1 << d)) - 1))
        O.apply_bcs(g, u, 1 << d, LDC_U[d])
        U.append(u)
    return U


def interior_div(g, U):
    return O.div(g, U[0], U[1], U[2], 0)[1:-1, 1:-1, 1:-1]


@pytest.mark.parametrize("alpha", [0.0, 1.0])
def test_projection_leaves_divergence_free_velocity(alpha):
    g = uniform_grid(10)
    U = noise_U(g, 1)
    st = S.MHDState(g, U, LDC_U, P_BCS, Re=400.0, dt=1e-3, alpha_U=alpha, isp_U=(200, 1e-12, 1e-15, 1),
                    isp_p=(2000, 1e-12, 1e-15, 1))
    for _ in range(2):
        st.step()
    assert np.abs(interior_div(g, st.U)).max() < 1e-9 * max(1.0, np.abs(st.U[0]).max() * 10)
    # the lid drives the flow along +x just below it
    assert st.U[0][3:-3, -2, 3:-3].mean() > 1e-4


def test_ct_source_divergence_free_and_rest_state():
    g = O.Grid(O.uniform_nodes(0, 1, 8), O.roberts_both_nodes(0, 1, 9, 1.1), O.uniform_nodes(0, 1, 7))
    U = noise_U(g, 3, 1e-2)
    st = S.MHDState(g, U, LDC_U, P_BCS, Re=100.0, dt=1e-3, B=[g.zeros(1 << d) for d in range(3)], Bbcs=PV_B,
                    phibcs=PHI_BCS, Rem=1.0, Ha=10.0)
    st.momentum_sources()
    st.solve_momentum()
    st.induction_sources()
    dv = O.div(g, st.FB[0], st.FB[1], st.FB[2], 0)[1:-1, 1:-1, 1:-1]
    assert np.abs(dv).max() < 1e-12 * max(1.0, max(np.abs(f).max() for f in st.FB))
    # no flow, no induced field: everything stays zero
    g2 = uniform_grid(6)
    st2 = S.MHDState(g2, [g2.zeros(1 << d) for d in range(3)], [O.BCs.make(["Dirichlet"] * 6)] * 3, P_BCS,
                     Re=100.0, dt=1e-3, B=[g2.zeros(1 << d) for d in range(3)], Bbcs=PV_B, phibcs=PHI_BCS, Ha=5.0)
    st2.step()
    for f in st2.U + st2.B:
        assert not np.any(f)
