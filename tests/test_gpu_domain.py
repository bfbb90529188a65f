"""GPU parity of the multi-domain induction: the library's domain ranges,
embed, extract and sigma_inv against the restatement bit for bit, and the
MHD step with the induction on the fluid mesh extended by walls and vacuum
against the oracle's step (every state field within 1e-10 per step, equal
solver iterations)."""
import numpy as np
import pytest

from oracle import domain_oracle as D
from oracle import oracle as O
from oracle import step_oracle as S

pytestmark = pytest.mark.gpu

LDC = [(["Dirichlet"] * 6, [0, 0, 0, 1.0, 0, 0]), (["Dirichlet"] * 6, None), (["Dirichlet"] * 6, None)]


def rel_l2(a, b):
    den = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b)) / (den if den > 0 else 1.0)


def meshes(gpu, N=(10, 9, 8), n_wall=1, n_ext=3, kind="uniform"):
    """Fluid nodes, the wall (sigma) mesh and the induction mesh.  uniform:
    ext_uniform_IO by n_wall and n_ext cells; roberts: the mesh bank's pattern
    (code/handwritten/sim_params/mesh_bank.f90:222-229): ext_Roberts_B_IO of a
    3-cell wall of thickness 0.5, then ext_Roberts_near_IO of 3 vacuum cells
    of length 1, built by the library's own generator."""
    hn = [O.uniform_nodes(0, 1, N[0]), O.roberts_both_nodes(0, 1, N[1], 1.1), O.uniform_nodes(0, 1, N[2])]
    if kind == "uniform":
        wall = [D.ext_uniform(h, n_wall) for h in hn]
        ind = [D.ext_uniform(h, n_ext) for h in hn]
    else:
        wall = [gpu.nodes_ext(h, "B", 0.5, 3) for h in hn]
        ind = [gpu.nodes_ext(w, "near", 1.0, 3) for w in wall]
        for w, i in zip(wall, ind):  # the library's generator equals the restatement
            np.testing.assert_array_equal(i, D.ext(w, "near", 1.0, 3))
    return hn, wall, ind


def test_domain_embed_extract_and_sigma_match_the_restatement(gpu):
    print("domain ranges, embed, extract, sigma", flush=True)
    hn, wall, ind = meshes(gpu)
    m, mw, mi = gpu.mesh(*hn), gpu.mesh(*wall), gpu.mesh(*ind)
    g, gi = O.Grid(*hn), O.Grid(*ind)
    md, mdw = gpu.mesh_domain(m, mi), gpu.mesh_domain(mw, mi)
    dom, domw = D.Domain(hn, ind), D.Domain(wall, ind)
    for d in range(3):
        assert md.overlap(d, False) == dom.C[d] and md.overlap(d, True) == dom.N[d]
        assert mdw.overlap(d, False) == domw.C[d] and mdw.overlap(d, True) == domw.N[d]
    rng = np.random.default_rng(1)
    for dl in (0, 1, 2, 4, 3, 5, 6, 7):
        small = np.asfortranarray(rng.standard_normal(g.shape(dl)))
        big = np.asfortranarray(rng.standard_normal(gi.shape(dl)))
        a, b = gpu.SF(m, dl).upload(small), gpu.SF(mi, dl).upload(big)
        gpu.embed(b, a, md)
        np.testing.assert_array_equal(b.download(), D.embed(big, small, dl, dom))
        c = gpu.SF(m, dl).fill(-1.0)
        gpu.extract(c, b, md)
        np.testing.assert_array_equal(c.download(), D.extract(np.full(g.shape(dl), -1.0, order="F"),
                                                              D.embed(big, small, dl, dom), dl, dom))
    for dl in S.EDGE_DL:
        s = gpu.SF(mi, dl)
        gpu.set_sigma_inv(s, mdw, 0.1)
        np.testing.assert_array_equal(s.download(), D.sigma_inv(gi.shape(dl), dl, domw, 0.1))
    with pytest.raises(gpu.MoonsError, match="overlap"):
        gpu.mesh_domain(m, gpu.mesh(*[O.uniform_nodes(5.0, 6.0, 4)] * 3))


@pytest.mark.parametrize("kind", ["uniform", "roberts"])
def test_multi_domain_step_matches_oracle(gpu, kind):
    """Fluid 10 x 9 x 8 (Roberts in y) inside an induction mesh extended at
    every wall (uniform: 3 cells, a one-cell wall; roberts: a Roberts-clustered
    3-cell wall and 3 vacuum cells); sigma = 1 on the fluid and the wall, 0.1
    beyond; B = 0 Dirichlet at the far boundary, phi Dirichlet; method 6 for U
    and B."""
    print(f"multi-domain mhd step ({kind} extension)", flush=True)
    hn, wall, ind = meshes(gpu, kind=kind)
    g, gi = O.Grid(*hn), O.Grid(*ind)
    m, mw, mi = gpu.mesh(*hn), gpu.mesh(*wall), gpu.mesh(*ind)
    md, mdw = gpu.mesh_domain(m, mi), gpu.mesh_domain(mw, mi)
    dom, domw = D.Domain(hn, ind), D.Domain(wall, ind)
    U0 = []
    for d in range(3):
        u = np.asfortranarray(0.3 * (2 * np.random.default_rng(30 + d).random(g.shape(1 << d)) - 1))
        O.apply_bcs(g, u, 1 << d, O.BCs.make(*LDC[d]))
        U0.append(u)
    isp = (5, 1e-6, 1e-13, 1)
    Bb = (["Dirichlet"] * 6, None)
    Ug = [gpu.SF(m, 1 << d).upload(U0[d]).set_BCs(*LDC[d]) for d in range(3)]
    pg = gpu.SF(m, "CC").fill(0.0).set_BCs(["Neumann"] * 6, None, all_Neumann=True)
    Bg = [gpu.SF(mi, 1 << d).fill(0.0).set_BCs(*Bb) for d in range(3)]
    phig = gpu.SF(mi, "CC").fill(0.0).set_BCs(["Dirichlet"] * 6)
    sig = [gpu.SF(mi, dl) for dl in S.EDGE_DL]
    gpu.set_sigma_inv(sig, mdw, 0.1)
    kw = dict(Re=400.0, dt=1e-3, alpha_U=1.0, Rem=1.0, Ha=10.0, B0=(0.0, 1.0, 0.0))
    st = gpu.MHD_step(m, Ug, pg, B=Bg, phi=phig, isp_U=isp, isp_p=isp, isp_B=isp, isp_phi=isp, m_ind=mi, domain=md,
                      sigma_inv=sig, **kw)
    ost = S.MHDState(g, U0, [O.BCs.make(*u) for u in LDC], O.BCs.make(["Neumann"] * 6), kw["Re"], kw["dt"],
                     alpha_U=1.0, B=[gi.zeros(1 << d) for d in range(3)], Bbcs=[O.BCs.make(*Bb)] * 3,
                     phibcs=O.BCs.make(["Dirichlet"] * 6), Rem=1.0, Ha=10.0, B0=(0.0, 1.0, 0.0), isp_U=isp,
                     isp_p=isp, isp_B=isp, isp_phi=isp, gi=gi, dom=dom,
                     sigInv=[D.sigma_inv(gi.shape(dl), dl, domw, 0.1) for dl in S.EDGE_DL])
    for n in range(3):
        info = st.step(1)
        ost.step()
        it = info["iterations"]
        assert (it["U"], it["p"], it["B"], it["phi"]) == (ost.iters["U"], ost.iters["p"], ost.iters["B"],
                                                         ost.iters["phi"]), (n, it, ost.iters)
        for d in range(3):
            assert rel_l2(Ug[d].download(), ost.U[d]) <= 1e-10, (n, "U", d)
            assert rel_l2(Bg[d].download(), ost.B[d]) <= 1e-10, (n, "B", d)
            assert rel_l2(st.field("J")[d].download(), ost.J[d]) <= 1e-10, (n, "J", d)
            assert rel_l2(st.field("jCrossB")[d].download(), ost.jCrossB[d]) <= 1e-10, (n, "JxB", d)
            assert rel_l2(st.field("FB")[d].download(), ost.FB[d]) <= 1e-10, (n, "FB", d)
        assert rel_l2(pg.download()[1:-1, 1:-1, 1:-1], ost.p[1:-1, 1:-1, 1:-1]) <= 1e-10, (n, "p")
        assert rel_l2(phig.download(), ost.phi) <= 1e-10, (n, "phi")
    assert max(np.abs(b.download()).max() for b in Bg) > 1e-6  # the flow induced a field
