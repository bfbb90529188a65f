"""Pins the C oracle against the reference's own routines, compiled by flang
from their sources (oracle/_ref): grids, stencil coefficients, the pencil
kernels of stencils.f90, the raw ghost-plane BC kernels, the Laplacian and
curl-curl diagonals of the Jacobi preconditioners and the staggered
interpolations with their ghost extrapolation."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="reference subset (oracle/_ref) not built here")


@pytest.mark.parametrize("N", [4, 8, 17, 32, 128])
def test_uniform_grid_matches_reference(N):
    ref = O.ref_grid(0, 0.0, 1.0, N)
    ours = O.uniform_nodes(0.0, 1.0, N)
    np.testing.assert_array_equal(ours[1:-1], ref)
    # ghost nodes mirror the first/last spacing (coordinates_extend.f90:529-538)
    assert ours[0] == ref[0] - (ref[1] - ref[0])
    assert ours[-1] == ref[-1] + (ref[-1] - ref[-2])


@pytest.mark.parametrize("N,beta", [(8, 1.05), (32, 1.05), (128, 1.05), (64, 1.2), (16, 3.0)])
def test_roberts_both_matches_reference(N, beta):
    ref = O.ref_grid(1, -1.0, 1.0, N, beta)
    ours = O.roberts_both_nodes(-1.0, 1.0, N, beta)
    np.testing.assert_allclose(ours[1:-1], ref, rtol=0, atol=4e-16)


def _grid(N, beta=None):
    if beta is None:
        h = O.uniform_nodes(0.0, 1.0, N)
    else:
        h = O.roberts_both_nodes(0.0, 1.0, N, beta)
    return O.Grid(h, h, h)


@pytest.mark.parametrize("beta", [None, 1.05, 1.3])
def test_stencil_coefficients_match_reference(beta):
    g = _grid(12, beta)
    dhn, dhc = g.coord(0, 2), g.coord(0, 3)
    sn, sc = len(dhn) + 1, len(dhc) + 1
    hn, hc = g.coord(0, 0), g.coord(0, 1)
    # staggered CC->N from dhc, N->CC from dhn
    _, D, U = O.ref_stencil(0, dhc, sc)
    np.testing.assert_array_equal(g.coord(0, 4), D)
    np.testing.assert_array_equal(g.coord(0, 5), U)
    _, D, U = O.ref_stencil(1, dhn, sn)
    np.testing.assert_array_equal(g.coord(0, 6), D)
    np.testing.assert_array_equal(g.coord(0, 7), U)
    # collocated CC n=1,2 and N n=1,2
    for kind, base, dh, s in ((2, 10, dhc, sc), (3, 13, dhc, sc), (4, 16, dhn, sn), (5, 19, dhn, sn)):
        L, D, U = O.ref_stencil(kind, dh, s)
        np.testing.assert_array_equal(g.coord(0, base + 0), L)
        np.testing.assert_array_equal(g.coord(0, base + 1), D)
        np.testing.assert_array_equal(g.coord(0, base + 2), U)
    D, U = O.ref_interp(hc, hn)
    np.testing.assert_array_equal(g.coord(0, 8), D)
    np.testing.assert_array_equal(g.coord(0, 9), U)


@pytest.mark.parametrize("beta", [None, 1.05])
@pytest.mark.parametrize("gen", [1, 2, 3])
def test_del_pencils_match_reference_kernels(beta, gen):
    """orc_del along each direction equals the reference pencil kernels applied
    pencil by pencil (stencils.f90 via ops_del.f90 diff_stag/diff_col)."""
    rng = np.random.default_rng(7)
    g = O.Grid(O.uniform_nodes(0, 1, 6) if beta is None else O.roberts_both_nodes(0, 1, 6, beta),
               O.roberts_both_nodes(0, 2, 7, 1.1), O.uniform_nodes(-1, 1, 5))
    for d in range(3):
        for dl_in, dl_out, n in ((0, 1 << d, 1), (1 << d, 0, 1), (0, 0, 1), (0, 0, 2), (1 << d, 1 << d, 1),
                                 (1 << d, 1 << d, 2)):
            f = np.asfortranarray(rng.standard_normal(g.shape(dl_in)))
            out0 = np.asfortranarray(rng.standard_normal(g.shape(dl_out)))
            ours = out0.copy(order="F")
            O.del_op(g, ours, dl_out, f, dl_in, n, d, 0, gen)
            nin, nout = (dl_in >> d) & 1, (dl_out >> d) & 1
            if not nin and nout:
                kind, gt, L, D, U = 0, 1, None, g.coord(d, 4), g.coord(d, 5)
            elif nin and not nout:
                kind, gt, L, D, U = 0, 0, None, g.coord(d, 6), g.coord(d, 7)
            elif not nin:
                kind, gt = 3, 0
                L, D, U = (g.coord(d, 10 + 3 * (n - 1) + t) for t in range(3))
            else:
                kind, gt = 6, 0
                L, D, U = (g.coord(d, 16 + 3 * (n - 1) + t) for t in range(3))
            kind += gen - 1
            ref = out0.copy(order="F")
            fm = np.moveaxis(f, d, -1)
            rm = np.moveaxis(ref, d, -1)
            for idx in np.ndindex(fm.shape[:-1]):
                rm[idx] = O.ref_pencil(kind, rm[idx], fm[idx], L if L is not None else D * 0, D, U, gt)
            np.testing.assert_array_equal(ours, ref)


def test_bc_rules_match_reference_kernels():
    """Ghost rules of orc_apply_bcs on each face equal the reference raw kernels."""
    rng = np.random.default_rng(3)
    h = O.roberts_both_nodes(0, 1, 6, 1.2)
    g = O.Grid(h, O.uniform_nodes(0, 2, 5), O.uniform_nodes(0, 1, 7))
    for dl in (0, 1, 2, 4, 7):
        for face in range(6):
            d = face // 2
            nal = (dl >> d) & 1
            for t, implicit, kind in (("Dirichlet", False, 1 if nal else 0), ("Neumann", False, 3 if nal else 2),
                                      ("Dirichlet", True, 6 if nal else 5), ("Neumann", True, 7)):
                types = ["none"] * 6
                types[face] = t
                vals = [0.0] * 6
                vals[face] = 0.37
                bcs = O.BCs.make(types, vals)
                u = np.asfortranarray(rng.standard_normal(g.shape(dl)))
                ours = u.copy(order="F")
                O.apply_bcs(g, ours, dl, bcs, implicit)
                s = u.shape[d]
                is_max = face % 2
                G = s - 1 if is_max else 0
                B = (s - 2 if is_max else 1) if nal else None
                I = (s - 3 if is_max else 2) if nal else (s - 2 if is_max else 1)
                um = np.moveaxis(u, d, 0)
                dh = O.ref_face_dh(g.coord(d, 0), face + 1)
                nhat = 1.0 if is_max else -1.0
                bv = np.full(um.shape[1:], 0.37)
                ub = um[B] if B is not None else np.zeros(um.shape[1:])
                ug, ubn = O.ref_bc_plane(kind, um[G], ub, um[I], bv, dh * nhat)
                exp = um.copy()
                exp[G] = ug
                if B is not None and t == "Dirichlet":
                    exp[B] = ubn
                np.testing.assert_array_equal(np.moveaxis(ours, d, 0), exp)


def test_face_dh_matches_reference_face_grid():
    """The dh every face rule uses is the reference's FSD%dh: dhn(1) of the
    face grid get_face_GI keeps (the last interior cell at a max face, not
    the ghost spacing)."""
    for hn in (O.roberts_both_nodes(0, 1, 9, 1.05), O.uniform_nodes(0, 1, 7), O.roberts_both_nodes(-1, 2, 12, 1.2),
               O.uniform_nodes(0, 3, 1)):
        g = O.Grid(hn, hn, hn)
        for face in range(6):
            assert O.face_dh(g, face // 2, face % 2) == O.ref_face_dh(hn, face + 1)


def test_robin_rules_match_reference_kernels():
    """Robin ghost rules (u + c_w du/dh = 0), direct and implicit, C and N
    located, equal the reference's raw kernels on every face, with the face
    sub-domain's dh from the reference's own face grid."""
    rng = np.random.default_rng(7)
    h = O.roberts_both_nodes(0, 1, 6, 1.2)
    g = O.Grid(h, O.uniform_nodes(0, 2, 5), O.roberts_both_nodes(-1, 1, 7, 1.05))
    cws = [0.01, 0.3, 1.0, 2.5, 0.05, 7.0]
    for dl in (0, 1, 2, 4, 3, 5, 6, 7):
        for face in range(6):
            d = face // 2
            nal = (dl >> d) & 1
            for implicit in (False, True):
                types = ["none"] * 6
                types[face] = "Robin"
                vals = [0.0] * 6
                vals[face] = 0.37
                bcs = O.BCs.make(types, vals, cws)
                u = np.asfortranarray(rng.standard_normal(g.shape(dl)))
                ours = u.copy(order="F")
                O.apply_bcs(g, ours, dl, bcs, implicit)
                s = u.shape[d]
                is_max = face % 2
                G = s - 1 if is_max else 0
                B = (s - 2 if is_max else 1) if nal else None
                I = (s - 3 if is_max else 2) if nal else (s - 2 if is_max else 1)
                um = np.moveaxis(u, d, 0)
                dh = O.ref_face_dh(g.coord(d, 0), face + 1)
                nhat = 1.0 if is_max else -1.0
                bv = np.full(um.shape[1:], 0.37)
                ub = um[B] if B is not None else np.zeros(um.shape[1:])
                kind = (1 if nal else 0) + (2 if implicit else 0)
                ug = O.ref_robin_plane(kind, um[G], ub, um[I], bv, dh, nhat, cws[face])
                exp = um.copy()
                exp[G] = ug
                np.testing.assert_array_equal(np.moveaxis(ours, d, 0), exp)
                # the rule is not the identity or a copy
                assert not np.array_equal(ug, um[G]) and not np.array_equal(ug, um[I])


DLS = {"CC": 0, "Node": 7, "Face_x": 1, "Face_y": 2, "Face_z": 4, "Edge_x": 6, "Edge_y": 5, "Edge_z": 3}


def _mixed_grid():
    return O.Grid(O.uniform_nodes(0.0, 1.0, 7), O.roberts_both_nodes(0.0, 1.0, 6, 1.2),
                  O.roberts_both_nodes(-1.0, 1.0, 8, 1.05))


@pytest.mark.parametrize("dl", list(DLS), ids=list(DLS))
def test_laplacian_diagonal_matches_reference(dl):
    """laplacian_diagonal (GF_diagonals.f90:15-35) bit for bit, wall rows and
    ghosts included: the Jacobi preconditioner's wall-row quirk (the diagonal
    ignores the BC-modified rows) is the reference's own."""
    g = _mixed_grid()
    np.testing.assert_array_equal(O.lap_diagonal(g, DLS[dl]), O.ref_laplacian_diagonal(g, DLS[dl]))


def test_curl_curl_diagonal_matches_reference():
    g = _mixed_grid()
    rng = np.random.default_rng(3)
    sig = [0.5 + rng.random(g.shape(e)) for e in (6, 5, 3)]
    ours = O.curl_curl_diagonal(g, sig)
    ref = O.ref_curl_curl_diagonal(g, sig)
    for d in range(3):
        np.testing.assert_array_equal(ours[d], ref[d])


def _pairs():
    out = []
    for name, dl in DLS.items():
        for d in range(3):
            out.append((name, dl, dl ^ (1 << d), d))
    return out


@pytest.mark.parametrize("name,dl_in,dl_out,d", _pairs(), ids=[f"{p[0]}-d{p[3]}" for p in _pairs()])
def test_interp_matches_reference(name, dl_in, dl_out, d):
    """interp (C2N with extrap_staggered at both ends, or N2C; GF_interp.f90:19-76,
    GF_extrap.f90:17-44) bit for bit for every location pair and direction on
    a mixed uniform/Roberts grid."""
    g = _mixed_grid()
    f = np.asfortranarray(np.random.default_rng(5 + d).standard_normal(g.shape(dl_in)))
    np.testing.assert_array_equal(O.interp(g, f, dl_in, dl_out, d), O.ref_interp_field(g, f, dl_in, dl_out, d))
