"""Pins the oracle's interpolation and constrained-transport operators:
interpolation is exact on linear fields (including the extrapolated ghost
nodes), and curl(E) of any edge field is divergence-free to round-off, so a
CT update B += dt*advect_B leaves div B unchanged."""
import numpy as np
import pytest

from oracle import oracle as O


def grid():
    return O.Grid(O.roberts_both_nodes(0, 1, 7, 1.2), O.uniform_nodes(0, 2, 6), O.roberts_both_nodes(-1, 1, 8, 1.05))


def coords(g, dl):
    return np.meshgrid(*[g.coord(d, 0) if (dl >> d) & 1 else g.coord(d, 1) for d in range(3)], indexing="ij")


@pytest.mark.parametrize("d", [0, 1, 2])
@pytest.mark.parametrize("dl_cc", [0, 1, 2, 4, 6, 5, 3])
def test_interp_exact_on_linear(d, dl_cc):
    g = grid()
    if (dl_cc >> d) & 1:
        return
    dl_n = dl_cc | (1 << d)
    X = coords(g, dl_cc)
    f = np.asfortranarray(2.0 * X[0] - 3.0 * X[1] + 0.5 * X[2] + 1.0)
    n = O.interp(g, f, dl_cc, dl_n, d)
    Xn = coords(g, dl_n)
    np.testing.assert_allclose(n, 2.0 * Xn[0] - 3.0 * Xn[1] + 0.5 * Xn[2] + 1.0, rtol=0, atol=1e-12)
    back = O.interp(g, np.asfortranarray(2.0 * Xn[0] - 3.0 * Xn[1] + 0.5 * Xn[2] + 1.0), dl_n, dl_cc, d)
    np.testing.assert_allclose(back, f, rtol=0, atol=1e-12)


def test_cross_product_against_numpy_on_constant_fields():
    g = grid()
    u = np.array([0.3, -1.2, 0.7])
    b = np.array([1.5, 0.25, -2.0])
    U_E = [np.full(g.shape(dl), u[c], order="F") for dl, c in ((6, 1), (6, 2), (5, 0), (5, 2), (3, 0), (3, 1))]
    B = [np.full(g.shape(1 << d), b[d], order="F") for d in range(3)]
    E = O.edge_cross_face_E(g, U_E, B)
    ref = np.cross(u, b)
    for d in range(3):
        np.testing.assert_allclose(E[d], ref[d], rtol=1e-14, atol=1e-14)


def test_advect_B_is_divergence_free():
    rng = np.random.default_rng(5)
    g = grid()
    U_E = [np.asfortranarray(rng.standard_normal(g.shape(dl))) for dl in (6, 6, 5, 5, 3, 3)]
    B = [np.asfortranarray(rng.standard_normal(g.shape(1 << d))) for d in range(3)]
    adv = O.advect_B(g, U_E, B)
    dv = O.div(g, adv[0], adv[1], adv[2], 0)
    scale = max(np.abs(a).max() for a in adv)
    assert np.abs(dv[1:-1, 1:-1, 1:-1]).max() < 1e-10 * scale
    # a CT update B + dt*adv keeps div B unchanged to round-off
    dt = 1e-3
    B2 = [np.asfortranarray(B[d] + dt * adv[d]) for d in range(3)]
    d0 = O.div(g, *B, 0)
    d1 = O.div(g, *B2, 0)
    assert np.abs((d1 - d0)[1:-1, 1:-1, 1:-1]).max() < 1e-12 * max(np.abs(b).max() for b in B) / g.coord(1, 2).min()
