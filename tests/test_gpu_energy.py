"""GPU parity of the energy equation with the oracle: the nrg_diffusion PCG
(iterations equal, solution within 1e-10), the energy step for each
time-marching method (1-5, the 3-stage RK included) with the reference's
zero k and with a conductivity, the sources (advection, diffusion, KE
diffusion, viscous dissipation, Joule and volumetric heating) bit for bit,
and the energy step coupled to the MHD step in solve_MOONS order (energy
first, on the current U)."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import step_oracle as S

pytestmark = pytest.mark.gpu

T_BCS = (["Dirichlet", "Dirichlet", "Neumann", "Neumann", "Neumann", "Neumann"], [1.0, 0.0, 0.0, 0.0, 0.0, 0.0])
LDC = [(["Dirichlet"] * 6, [0, 0, 0, 1.0, 0, 0]), (["Dirichlet"] * 6, None), (["Dirichlet"] * 6, None)]


def rel_l2(a, b):
    den = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b)) / (den if den > 0 else 1.0)


def mesh(gpu, N=(10, 9, 8)):
    hs = [O.uniform_nodes(0, 1, N[0]), O.roberts_both_nodes(0, 1, N[1], 1.1), O.uniform_nodes(0, 1, N[2])]
    return O.Grid(*hs), gpu.mesh(*hs)


def face_fields(g, fn):
    return [np.asfortranarray(fn(d, g.shape(1 << d))) for d in range(3)]


def test_nrg_pcg_matches_oracle(gpu):
    print("nrg pcg", flush=True)
    g, m = mesh(gpu)
    rng = np.random.default_rng(0)
    k = face_fields(g, lambda d, s: 0.5 + rng.random(s))
    bcs = (["Neumann", "Dirichlet", "Neumann", "Neumann", "Dirichlet", "Neumann"], [0.3, -1.0, 0.2, 0.0, 2.0, -0.5])
    b = np.zeros(g.shape(0), order="F")
    b[1:-1, 1:-1, 1:-1] = rng.standard_normal(tuple(n - 2 for n in g.shape(0)))
    c = -0.05
    xo = np.zeros(g.shape(0), order="F")
    it, res0, res, _ = O.pcg_nrg(g, O.BCs.make(*bcs), False, k, c, xo, b, 200, 1e-10, 1e-13)
    x = gpu.SF(m, "CC").fill(0.0).set_BCs(*bcs)
    P = gpu.PCG_solver_SF(m, x, gpu.iter_solver_params(200, 1e-10, 1e-13, 1), op="nrg_diffusion")
    P.set_k([gpu.SF(m, 1 << d).upload(k[d]) for d in range(3)])
    P.update_MFP(c)
    info = P.solve(x, gpu.SF(m, "CC").upload(b))
    assert info.iterations == it, (info, it)
    assert abs(info.res_norm0 - res0) <= 1e-12 * res0
    assert rel_l2(x.download(), xo) <= 1e-10


def build(gpu, method, k_kind, N=(10, 9, 8)):
    g, m = mesh(gpu, N)
    X = np.meshgrid(g.coord(0, 1), g.coord(1, 1), g.coord(2, 1), indexing="ij")[0]
    T0 = np.asfortranarray(1.0 - X + 0.05 * np.sin(7.0 * X))
    O.apply_bcs(g, T0, 0, O.BCs.make(*T_BCS))
    U0 = []
    for d in range(3):
        u = np.asfortranarray(0.3 * (2 * np.random.default_rng(20 + d).random(g.shape(1 << d)) - 1))
        O.apply_bcs(g, u, 1 << d, O.BCs.make(*LDC[d]))
        U0.append(u)
    if k_kind == "zero":
        k = face_fields(g, lambda d, s: np.zeros(s))
    else:
        k = face_fields(g, lambda d, s: np.ones(s) * (1.0 + 0.5 * d))
    Pe, dt = 50.0, 2e-3
    explicit_diff = method != 3
    kw = dict(dt=dt, method=method, advection=(True, -1.0), diffusion=(explicit_diff, 1.0 / Pe))
    isp = (50, 1e-12, 1e-15, 1)
    T = gpu.SF(m, "CC").upload(T0).set_BCs(*T_BCS)
    U = [gpu.SF(m, 1 << d).upload(U0[d]).set_BCs(*LDC[d]) for d in range(3)]
    kg = [gpu.SF(m, 1 << d).upload(k[d]) for d in range(3)]
    st = gpu.Energy_step(m, T, U, kg, Pe=Pe, isp_T=isp, **kw)
    ost = S.EnergyState(g, T0, O.BCs.make(*T_BCS), k, coeff_implicit=-1.0 / Pe, isp=isp, **kw)
    return g, m, st, ost, T, U, U0


@pytest.mark.parametrize("method,k_kind", [(1, "zero"), (2, "zero"), (3, "zero"), (3, "conductivity")])
def test_energy_step_matches_oracle(gpu, method, k_kind):
    print(f"energy method {method} k={k_kind}", flush=True)
    g, m, st, ost, T, U, U0 = build(gpu, method, k_kind)
    for n in range(3):
        info = st.step(1)
        ost.step(U0)
        if n == 0:  # the first sources are element-wise and stencil work only
            np.testing.assert_array_equal(st.field("F").download(), ost.F)
        assert rel_l2(st.field("F").download(), ost.F) <= 1e-10, (n, "F")
        assert rel_l2(T.download(), ost.T) <= 1e-10, (n, "T")
        if method == 3:
            assert info["iterations"] == ost.iters, (n, info, ost.iters)
            assert rel_l2(st.field("Tnm1").download(), ost.Tnm1) <= 1e-10
        elif n == 0:  # explicit methods: the first update is element-wise too
            np.testing.assert_array_equal(T.download(), ost.T)


def test_energy_coupled_with_the_mhd_step(gpu):
    """solve_MOONS order: energy (on the current U), then momentum."""
    from test_gpu_step import build as build_mhd

    print("energy + mhd", flush=True)
    isp = (5, 1e-6, 1e-13, 1)
    N = (10, 9, 8)
    g, m, mst, most, Ug, pg, kw = build_mhd(gpu, N, False, 1.0, isp)
    X = np.meshgrid(g.coord(0, 1), g.coord(1, 1), g.coord(2, 1), indexing="ij")[0]
    T0 = np.asfortranarray(1.0 - X)
    O.apply_bcs(g, T0, 0, O.BCs.make(*T_BCS))
    k = face_fields(g, lambda d, s: np.ones(s))
    nkw = dict(dt=1e-3, method=3, advection=(True, -1.0), diffusion=(False, 0.0))
    T = gpu.SF(m, "CC").upload(T0).set_BCs(*T_BCS)
    est = gpu.Energy_step(m, T, Ug, [gpu.SF(m, 1 << d).upload(k[d]) for d in range(3)], Pe=100.0,
                          isp_T=(50, 1e-12, 1e-15, 1), **nkw)
    eost = S.EnergyState(g, T0, O.BCs.make(*T_BCS), k, coeff_implicit=-0.01, isp=(50, 1e-12, 1e-15, 1), **nkw)
    for n in range(2):
        est.step(1)
        eost.step(most.U)
        mst.step(1)
        most.step()
        assert rel_l2(T.download(), eost.T) <= 1e-10, (n, "T")
        for d in range(3):
            assert rel_l2(Ug[d].download(), most.U[d]) <= 1e-10, (n, "U", d)


EDGE_DL = (6, 5, 3)


def build_sources(gpu, method, k_kind, sources, rk_stages=1, diffusion=True, N=(10, 9, 8)):
    """The energy step with the given extra sources ({name: scale}) on random
    U, J and sigmaInv_CC; returns the device and oracle states."""
    g, m = mesh(gpu, N)
    rng = np.random.default_rng(7)
    X = np.meshgrid(g.coord(0, 1), g.coord(1, 1), g.coord(2, 1), indexing="ij")[0]
    T0 = np.asfortranarray(1.0 - X + 0.05 * np.sin(7.0 * X))
    O.apply_bcs(g, T0, 0, O.BCs.make(*T_BCS))
    U0 = []
    for d in range(3):
        u = np.asfortranarray(0.3 * (2 * rng.random(g.shape(1 << d)) - 1))
        O.apply_bcs(g, u, 1 << d, O.BCs.make(*LDC[d]))
        U0.append(u)
    J0 = [np.asfortranarray(rng.standard_normal(g.shape(EDGE_DL[d]))) for d in range(3)]
    sig0 = np.asfortranarray(0.5 + rng.random(g.shape(0)))
    if k_kind == "zero":
        k = face_fields(g, lambda d, s: np.zeros(s))
    else:
        k = face_fields(g, lambda d, s: np.ones(s) * (1.0 + 0.5 * d))
    Pe, dt = 50.0, 2e-3
    kw = dict(dt=dt, method=method, advection=(True, -1.0), diffusion=(diffusion, 1.0 / Pe), rk_stages=rk_stages)
    for name in ("KE_diffusion", "viscous_dissipation", "joule_heating", "volumetric_heating"):
        kw[name] = (name in sources, sources.get(name, 1.0))
    isp = (50, 1e-12, 1e-15, 1)
    T = gpu.SF(m, "CC").upload(T0).set_BCs(*T_BCS)
    U = [gpu.SF(m, 1 << d).upload(U0[d]).set_BCs(*LDC[d]) for d in range(3)]
    J = [gpu.SF(m, EDGE_DL[d]).upload(J0[d]) for d in range(3)]
    sig = gpu.SF(m, "CC").upload(sig0)
    kg = [gpu.SF(m, 1 << d).upload(k[d]) for d in range(3)]
    st = gpu.Energy_step(m, T, U, kg, Pe=Pe, isp_T=isp, J=J, sigma_inv=sig, **kw)
    ost = S.EnergyState(g, T0, O.BCs.make(*T_BCS), k, coeff_implicit=-1.0 / Pe, isp=isp, J=J0, sigma_inv_cc=sig0,
                        **kw)
    return st, ost, T, U0, (J, sig)


SOURCE_SETS = [
    {"KE_diffusion": 0.02},
    {"viscous_dissipation": 0.5},
    {"joule_heating": 0.25},
    {"volumetric_heating": 0.7},
    {"KE_diffusion": 0.02, "viscous_dissipation": 0.5, "joule_heating": 0.25, "volumetric_heating": 0.7},
]


@pytest.mark.parametrize("sources", SOURCE_SETS, ids=lambda s: "+".join(s))
def test_energy_sources_bit_exact(gpu, sources):
    """KE diffusion, viscous dissipation, Joule and volumetric heating: the
    sources F bit for bit at every step (explicit Euler keeps T element-wise
    too), T bit for bit."""
    print(f"energy sources {sorted(sources)}", flush=True)
    st, ost, T, U0, _ = build_sources(gpu, 1, "zero", sources)
    for n in range(3):
        st.step(1)
        ost.step(U0)
        np.testing.assert_array_equal(st.field("F").download(), ost.F)
        np.testing.assert_array_equal(T.download(), ost.T)


@pytest.mark.parametrize("method,rk_stages,k_kind", [(4, 1, "zero"), (4, 1, "conductivity"), (5, 1, "conductivity"),
                                                      (5, 3, "zero"), (5, 3, "conductivity"), (3, 3, "conductivity")])
def test_energy_methods_4_5_match_oracle(gpu, method, rk_stages, k_kind):
    """BDF2 (method 4) and the RK stage's Euler (method 5, RK inactive or the
    3-stage RK) with every source on: T, Tnm1, F and L within 1e-10 and the
    solve's iterations equal at each step."""
    print(f"energy method {method} rk {rk_stages} k={k_kind}", flush=True)
    srcs = SOURCE_SETS[-1]
    st, ost, T, U0, _ = build_sources(gpu, method, k_kind, srcs, rk_stages=rk_stages)
    for n in range(3):
        info = st.step(1)
        ost.step(U0)
        assert rel_l2(st.field("F").download(), ost.F) <= 1e-10, (n, "F")
        assert rel_l2(st.field("L").download(), ost.L) <= 1e-10, (n, "L")
        assert rel_l2(T.download(), ost.T) <= 1e-10, (n, "T")
        assert rel_l2(st.field("Tnm1").download(), ost.Tnm1) <= 1e-10, (n, "Tnm1")
        assert info["iterations"] == ost.iters, (n, info, ost.iters)


def test_energy_joule_needs_its_fields(gpu):
    g, m = mesh(gpu)
    T = gpu.SF(m, "CC").fill(0.0).set_BCs(*T_BCS)
    st = gpu.Energy_step(m, T, None, None, method=1, advection=(False, 0.0), joule_heating=(True, 1.0))
    with pytest.raises(gpu.MoonsError, match="Joule"):
        st.step(1)
    with pytest.raises(gpu.MoonsError, match="edge"):
        st.set_joule([gpu.SF(m, "CC").fill(0.0)] * 3, gpu.SF(m, "CC").fill(1.0))
