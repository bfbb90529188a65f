"""Pins the multi-domain restatement (oracle/domain_oracle.py) and the
library's host-side mesh extension: the uniform extension and the overlap
against the reference's own ext_uniform_IO and get_C/N_overlap (oracle/_ref,
where built), and by what the geometry fixes -- the extension keeps every
original node and continues each end's ghost spacing, the overlap pairs
coinciding points and covers the whole interior mesh, embed and extract
are inverse on the overlap, and the conductivity is 1 exactly on the
conducting sub-mesh."""
import numpy as np
import pytest

import moons_amd as M
from oracle import domain_oracle as D
from oracle import oracle as O
from oracle import step_oracle as S

NODES = [O.uniform_nodes(0.0, 1.0, 6), O.roberts_both_nodes(-1.0, 1.0, 9, 1.1), O.uniform_nodes(-2.0, 0.5, 4)]
ref_only = pytest.mark.skipif(not O.ref_available(), reason="reference subset (oracle/_ref) not built here")


@pytest.mark.parametrize("N", [1, 4, 11])
@pytest.mark.parametrize("which", range(3))
def test_extension_keeps_the_mesh_and_continues_the_end_spacing(which, N):
    hn = NODES[which]
    ext = D.ext_uniform(hn, N)
    assert ext.size == hn.size + 2 * N
    np.testing.assert_array_equal(ext[N + 1:N + hn.size - 1], hn[1:-1])  # every non-ghost node kept
    d0, d1 = hn[1] - hn[0], hn[-1] - hn[-2]  # the ghost spacings
    np.testing.assert_allclose(np.diff(ext[:N + 2]), d0, rtol=1e-12)
    np.testing.assert_allclose(np.diff(ext[-N - 2:]), d1, rtol=1e-12)
    # the library's host restatement is the same
    np.testing.assert_array_equal(M.nodes_ext_uniform(hn, N), ext)


@ref_only
@pytest.mark.parametrize("N", [1, 4, 11])
@pytest.mark.parametrize("which", range(3))
def test_extension_equals_the_reference(which, N):
    np.testing.assert_array_equal(D.ext_uniform(NODES[which], N), O.ref_ext_uniform(NODES[which], N))


def centres(hn):
    hn = np.asarray(hn)
    return (hn[1:] + hn[:-1]) / 2.0


def pairs():
    a = NODES[1]
    return [
        (a, D.ext_uniform(a, 3)),           # fluid inside its extension
        (a, a),                              # the same mesh
        (D.ext_uniform(a, 1), D.ext_uniform(a, 4)),
        (O.uniform_nodes(0.0, 1.0, 8), O.uniform_nodes(0.5, 1.5, 8)),  # half overlapping
    ]


@pytest.mark.parametrize("case", range(4))
def test_overlap_pairs_coinciding_points(case):
    a, b = pairs()[case]
    for h1, h2 in ((centres(a), centres(b)), (a, b)):
        r = D.overlap(h1, h2)
        assert r is not None and r[1] - r[0] == r[3] - r[2]
        np.testing.assert_allclose(h1[r[0]:r[1] + 1], h2[r[2]:r[3] + 1], rtol=0, atol=1e-12)
        # nothing coincides outside the ranges (ghost points left out)
        for i in range(1, len(h1) - 1):
            hit = np.any(np.abs(h2[1:-1] - h1[i]) < 1e-12)
            assert hit == (r[0] <= i <= r[1])


@ref_only
@pytest.mark.parametrize("case", range(4))
def test_overlap_equals_the_reference(case):
    a, b = pairs()[case]
    rc, rn = D.overlap(centres(a), centres(b)), D.overlap(a, b)
    assert O.ref_overlap("C", a, b) == tuple(v + 1 for v in rc)
    assert O.ref_overlap("N", a, b) == tuple(v + 1 for v in rn)


def test_fluid_domain_covers_every_interior_cell_and_node():
    hn = [NODES[0], NODES[1], NODES[2]]
    ext = [D.ext_uniform(h, 3) for h in hn]
    dom = D.Domain(hn, ext)
    for d in range(3):
        sn = len(hn[d])
        assert dom.C[d][:2] == (1, sn - 3) and dom.N[d][:2] == (1, sn - 2)
        assert dom.C[d][2] == 1 + 3 and dom.N[d][2] == 1 + 3


@pytest.mark.parametrize("dl", [0, 1, 2, 4, 3, 5, 6, 7])
def test_embed_and_extract_are_inverse_on_the_overlap(dl):
    hn = [NODES[0], NODES[1], NODES[2]]
    ext = [D.ext_uniform(h, 2) for h in hn]
    g, gt = O.Grid(*hn), O.Grid(*ext)
    dom = D.Domain(hn, ext)
    rng = np.random.default_rng(dl)
    small = np.asfortranarray(rng.standard_normal(g.shape(dl)))
    big = np.asfortranarray(rng.standard_normal(gt.shape(dl)))
    emb = D.embed(big, small, dl, dom)
    si, st = dom.ranges(dl)
    np.testing.assert_array_equal(emb[st], small[si])
    mask = np.ones(big.shape, bool)
    mask[st] = False
    np.testing.assert_array_equal(emb[mask], big[mask])
    back = D.extract(np.zeros(g.shape(dl), order="F"), emb, dl, dom)
    np.testing.assert_array_equal(back[si], small[si])


def test_sigma_inv_is_one_exactly_on_the_conducting_mesh():
    hn = [NODES[0], NODES[1], NODES[2]]
    wall = [D.ext_uniform(h, 1) for h in hn]
    ind = [D.ext_uniform(h, 3) for h in hn]
    gi = O.Grid(*ind)
    dom = D.Domain(wall, ind)
    for dl in S.EDGE_DL:
        s = D.sigma_inv(gi.shape(dl), dl, dom, 1e-2)
        _, st = dom.ranges(dl)
        assert np.all(s[st] == 1.0)
        mask = np.ones(s.shape, bool)
        mask[st] = False
        assert np.all(s[mask] == 100.0)


def test_extended_ct_source_is_divergence_free_and_vanishes_in_the_vacuum():
    """The CT source curl(U_E x B0) on the extended mesh, with the fluid's edge
    velocity embedded (zero outside the fluid): its divergence is at round-off
    everywhere, and it is exactly zero on faces more than one cell away from
    the fluid (E = 0 on every edge around them)."""
    N = (6, 5, 4)
    hn = [O.uniform_nodes(0, 1, N[0]), O.roberts_both_nodes(0, 1, N[1], 1.1), O.uniform_nodes(0, 1, N[2])]
    ext = [D.ext_uniform(h, 3) for h in hn]
    g, gi = O.Grid(*hn), O.Grid(*ext)
    dom = D.Domain(hn, ext)
    wall = D.Domain([D.ext_uniform(h, 1) for h in hn], ext)
    sig = [D.sigma_inv(gi.shape(dl), dl, wall, 0.1) for dl in S.EDGE_DL]
    ldc = [O.BCs.make(["Dirichlet"] * 6, [0, 0, 0, 1.0, 0, 0]), O.BCs.make(["Dirichlet"] * 6),
           O.BCs.make(["Dirichlet"] * 6)]
    U = []
    for d in range(3):
        u = np.asfortranarray(0.3 * (2 * np.random.default_rng(d).random(g.shape(1 << d)) - 1))
        O.apply_bcs(g, u, 1 << d, ldc[d])
        U.append(u)
    st = S.MHDState(g, U, ldc, O.BCs.make(["Neumann"] * 6), 400.0, 1e-2, B=[gi.zeros(1 << d) for d in range(3)],
                    Bbcs=[O.BCs.make(["Dirichlet"] * 6)] * 3, phibcs=O.BCs.make(["Dirichlet"] * 6), Ha=5.0,
                    gi=gi, dom=dom, sigInv=sig, B0=(0.2, 1.0, -0.3))
    st.step()
    FB = st.FB
    divF = O.div(gi, FB[0], FB[1], FB[2], 0)[1:-1, 1:-1, 1:-1]
    scale = max(np.abs(f).max() for f in FB) / min(np.diff(h).min() for h in ext)
    assert max(np.abs(f).max() for f in FB) > 1e-3
    assert np.abs(divF).max() <= 1e-12 * scale
    for d in range(3):
        _, st_rng = dom.ranges(1 << d)
        far = np.ones(FB[d].shape, bool)
        grown = tuple(slice(max(r.start - 2, 0), r.stop + 2) for r in st_rng)
        far[grown] = False
        assert np.all(FB[d][far] == 0.0)
