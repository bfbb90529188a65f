"""Pins the oracle's energy equation by what the mathematics fixes: the
nrg_diffusion operator against the pinned Laplacian (k = 1) and the identity
(k = 0), symmetry and positivity of its volume-weighted matrix, its PCG
against a dense solve, the energy sources on fields they differentiate
exactly (advection, diffusion, KE diffusion, viscous dissipation, Joule and
volumetric heating), the explicit methods' closed form with the AB2 start-up
(Fnm1 = 0), backward Euler on a discrete eigenmode, BDF2's and the 3-stage
RK's exact integration of a constant source, and the RK's explicit diffusion
weights."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import step_oracle as S


def grid(Ns=(5, 4, 6), betas=(None, 1.2, 1.05)):
    hs = [O.uniform_nodes(0.0, 1.0, N) if b is None else O.roberts_both_nodes(0.0, 1.0, N, b) for N, b in zip(Ns, betas)]
    return O.Grid(*hs)


def faces(g, fn):
    return [np.asfortranarray(fn(d, g.shape(1 << d))) for d in range(3)]


def varied_k(g, seed=0):
    rng = np.random.default_rng(seed)
    return faces(g, lambda d, s: 0.5 + rng.random(s))


def interior(a):
    return a[1:-1, 1:-1, 1:-1]


MIXED = ["Neumann", "Dirichlet", "Neumann", "Neumann", "Dirichlet", "Neumann"]


def test_unit_k_gives_lap_times_c_plus_x_bit_for_bit():
    g = grid()
    bcs = O.BCs.make(MIXED)
    x = np.asfortranarray(np.random.default_rng(1).standard_normal(g.shape(0)))
    k1 = faces(g, lambda d, s: np.ones(s))
    Ax = O.nrg_op(g, x.copy(order="F"), k1, -0.37, bcs)
    xi = x.copy(order="F")
    O.apply_bcs(g, xi, 0, bcs, implicit=True)
    ref = O.lap_centered(g, xi, 0) * (-0.37) + xi
    np.testing.assert_array_equal(interior(Ax), interior(ref))


def test_zero_k_is_the_identity():
    g = grid()
    x = np.asfortranarray(np.random.default_rng(2).standard_normal(g.shape(0)))
    Ax = O.nrg_op(g, x.copy(order="F"), faces(g, lambda d, s: np.zeros(s)), -5.0, O.BCs.make(MIXED))
    np.testing.assert_array_equal(interior(Ax), interior(x))


def assemble(g, bcs, k, c):
    s = g.shape(0)
    vol = O.volume(g, 0)
    idx = [(i, j, kk) for kk in range(1, s[2] - 1) for j in range(1, s[1] - 1) for i in range(1, s[0] - 1)]
    A = np.zeros((len(idx), len(idx)))
    for col, q in enumerate(idx):
        e = np.zeros(s, order="F")
        e[q] = 1.0
        Ae = O.nrg_op(g, e, k, c, bcs) * vol
        A[:, col] = [Ae[p] for p in idx]
    return A, idx


@pytest.mark.parametrize("kind", ["Neumann", "Dirichlet", "mixed"])
def test_volume_weighted_operator_is_symmetric_positive_and_linear_in_ck(kind):
    g = grid()
    types = {"Neumann": ["Neumann"] * 6, "Dirichlet": ["Dirichlet"] * 6, "mixed": MIXED}[kind]
    bcs = O.BCs.make(types)
    k = varied_k(g)
    A, _ = assemble(g, bcs, k, -0.05)
    np.testing.assert_allclose(A, A.T, rtol=0, atol=1e-12 * np.abs(A).max())
    assert np.linalg.eigvalsh(A)[0] > 0
    # (I + c div(k grad)) is affine in the product c k
    A2, _ = assemble(g, bcs, [2.0 * a for a in k], -0.025)
    np.testing.assert_allclose(A2, A, rtol=0, atol=1e-13 * np.abs(A).max())


def random_interior(g, seed):
    b = np.zeros(g.shape(0), order="F")
    b[1:-1, 1:-1, 1:-1] = np.random.default_rng(seed).standard_normal(interior(b).shape)
    return b


@pytest.mark.parametrize("kind", ["Neumann", "Dirichlet", "mixed"])
def test_pcg_matches_dense_solve(kind):
    g = grid()
    types = {"Neumann": ["Neumann"] * 6, "Dirichlet": ["Dirichlet"] * 6, "mixed": MIXED}[kind]
    bcs = O.BCs.make(types)
    k, c = varied_k(g, 3), -0.08
    b = random_interior(g, 4)
    x = np.zeros(g.shape(0), order="F")
    it, res0, res, _ = O.pcg_nrg(g, bcs, False, k, c, x, b, 500, 1e-14, 1e-300)
    assert res / res0 < 1e-13
    A, idx = assemble(g, bcs, k, c)
    vol = O.volume(g, 0)
    xd = np.linalg.solve(A, np.array([b[q] * vol[q] for q in idx]))
    np.testing.assert_allclose([x[q] for q in idx], xd, rtol=0, atol=1e-10 * np.abs(xd).max())


def test_pcg_honours_nonhomogeneous_bc_values():
    """With BC values the solution, its ghosts set from those values, satisfies
    the operator applied to the stored ghosts (compute_Ax_BC splits off the
    BC part of the right-hand side)."""
    g = grid()
    bcs = O.BCs.make(MIXED, [0.3, -1.0, 0.2, 0.0, 2.0, -0.5])
    k, c = varied_k(g, 5), -0.06
    b = random_interior(g, 6)
    x = np.zeros(g.shape(0), order="F")
    O.pcg_nrg(g, bcs, False, k, c, x, b, 500, 1e-14, 1e-300)
    Ax = O.nrg_op(g, x.copy(order="F"), k, c, O.BCs.make(["none"] * 6))
    np.testing.assert_allclose(interior(Ax), interior(b), rtol=0, atol=1e-10 * np.abs(b).max())


def cell_coords(g):
    return np.meshgrid(g.coord(0, 1), g.coord(1, 1), g.coord(2, 1), indexing="ij")


def test_diffusion_source_exact_on_quadratic_T():
    g = grid((6, 5, 7))
    X, Y, Z = cell_coords(g)
    T = np.asfortranarray(0.7 * X**2 - 1.3 * Y**2 + 0.4 * Z**2 + X * 0.1)
    st = S.EnergyState(g, T, O.BCs.make(["none"] * 6), faces(g, lambda d, s: np.zeros(s)), 1e-3, method=1,
                       advection=(False, 0.0), diffusion=(True, 0.25))
    st.sources(None)
    # cells next to the walls use the one-sided collocated rows (pad 0)
    deep = (slice(2, -2),) * 3
    np.testing.assert_allclose(st.F[deep], 0.25 * 2 * (0.7 - 1.3 + 0.4), rtol=1e-10)


def test_advection_source_exact_on_linear_T_and_uniform_U():
    g = grid((6, 5, 7))
    X, Y, Z = cell_coords(g)
    a = (0.6, -1.1, 0.35)
    T = np.asfortranarray(a[0] * X + a[1] * Y + a[2] * Z + 2.0)
    u = (0.9, -0.4, 1.7)
    U = faces(g, lambda d, s: np.full(s, u[d]))
    st = S.EnergyState(g, T, O.BCs.make(["none"] * 6), faces(g, lambda d, s: np.zeros(s)), 1e-3, method=1,
                       advection=(True, -1.0), diffusion=(False, 0.0))
    st.sources(U)
    # F = scale * (-div(T U)) = -1 * (-(u . grad T))
    np.testing.assert_allclose(interior(st.F), sum(u[d] * a[d] for d in range(3)), rtol=1e-11)


@pytest.mark.parametrize("method", [1, 2])
def test_explicit_methods_closed_form_with_ab2_start_up(method):
    """A quadratic T has a constant diffusion source s: Euler gives T0 + s dt n;
    AB2 starts from Fnm1 = 0 (the reference's start-up), so T0 + s dt (n + 1/2)
    after n >= 1 steps -- away from the walls, where the one-sided collocated
    rows of the cells next to the walls reach one more cell inward each step."""
    g = grid((10, 10, 10), (None, None, None))
    X, Y, Z = cell_coords(g)
    T0 = np.asfortranarray(X**2 + 0.5 * Y**2 - 0.25 * Z**2)
    s = 0.5 * 2 * (1.0 + 0.5 - 0.25)
    dt = 1e-2
    st = S.EnergyState(g, T0, O.BCs.make(["none"] * 6), faces(g, lambda d, sh: np.zeros(sh)), dt, method=method,
                       advection=(False, 0.0), diffusion=(True, 0.5))
    for n in (1, 2):
        st.step(None)
        expect = T0 + s * dt * (n if method == 1 else n + 0.5)
        deep = (slice(1 + n, -1 - n),) * 3
        np.testing.assert_allclose(st.T[deep], expect[deep], rtol=0, atol=1e-12)


def test_backward_euler_on_a_discrete_eigenmode():
    """cos(pi x) at the cell centres with homogeneous Neumann walls is an
    eigenvector of the staggered Laplacian, eigenvalue (2 cos(pi h) - 2)/h^2;
    one method-3 step with k = 1 and no sources divides it by
    1 + c lambda, c = dt coeff_implicit."""
    N = 12
    g = grid((N, 6, 5), (None, None, None))
    X, _, _ = cell_coords(g)
    T0 = np.asfortranarray(np.cos(np.pi * X))
    h = 1.0 / N
    lam = (2.0 * np.cos(np.pi * h) - 2.0) / h**2
    Pe, dt = 10.0, 0.05
    c = dt * (-1.0 / Pe)
    st = S.EnergyState(g, T0, O.BCs.make(["Neumann"] * 6), faces(g, lambda d, s: np.ones(s)), dt, method=3,
                       coeff_implicit=-1.0 / Pe, advection=(False, 0.0), diffusion=(False, 0.0),
                       isp=(200, 1e-14, 1e-300, 1))
    st.step(None)
    np.testing.assert_allclose(interior(st.T), interior(T0) / (1.0 + c * lam), rtol=0, atol=1e-11)
    np.testing.assert_array_equal(st.Tnm1, T0)
    # with k = 0 (the reference's binding) the solve returns its right-hand side
    st0 = S.EnergyState(g, T0, O.BCs.make(["Neumann"] * 6), faces(g, lambda d, s: np.zeros(s)), dt, method=3,
                        coeff_implicit=-1.0 / Pe, advection=(False, 0.0), diffusion=(False, 0.0),
                        isp=(200, 1e-14, 1e-300, 1))
    st0.step(None)
    np.testing.assert_allclose(interior(st0.T), interior(T0), rtol=0, atol=1e-13)


# ---------------------------------------------------------------------------
# KE diffusion, viscous dissipation, Joule and volumetric heating; methods 4, 5
# ---------------------------------------------------------------------------

def ugrid(N=(8, 7, 9)):
    return O.Grid(*[O.uniform_nodes(0.0, 1.0, n) for n in N])


def coords(g, dl):
    """Per-direction coordinate arrays of a field at location dl (N: nodes, else centres)."""
    return [g.coord(d, 0) if (dl >> d) & 1 else g.coord(d, 1) for d in range(3)]


def field_of(g, dl, fn):
    x, y, z = coords(g, dl)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    return np.asfortranarray(fn(X, Y, Z))


def deep(a, w=2):
    return a[w:-w, w:-w, w:-w]


def test_cell_center_velocity_of_linear_faces_is_exact():
    g = grid()
    U = [field_of(g, 1 << d, lambda X, Y, Z, d=d: (1.0 + d) * (X, Y, Z)[d] + 0.5 * Z) for d in range(3)]
    U_CC = S.cell_center_velocity(g, U)
    for d in range(3):
        ref = field_of(g, 0, lambda X, Y, Z, d=d: (1.0 + d) * (X, Y, Z)[d] + 0.5 * Z)
        np.testing.assert_allclose(U_CC[d], ref, rtol=0, atol=1e-13)


def test_ke_diffusion_of_linear_velocity_is_constant():
    # U = (x, 2y, 3z) -> |U|^2 = x^2 + 4y^2 + 9z^2 -> lap = 2 + 8 + 18 = 28
    g = ugrid()
    U = [field_of(g, 1 << d, lambda X, Y, Z, d=d: (1.0 + d) * (X, Y, Z)[d]) for d in range(3)]
    t = S.ke_diffusion_term(g, S.cell_center_velocity(g, U))
    np.testing.assert_allclose(deep(t), 28.0, rtol=1e-9)


def test_viscous_dissipation_is_the_full_gradient_norm():
    # U_x = 3y + z, U_y = -2x + 5z, U_z = 4x: grad U : grad U = 9 + 1 + 4 + 25 + 16 = 55;
    # a diagonal-only or transposed-square mistake gives 0 or another value
    g = ugrid()
    fns = [lambda X, Y, Z: 3.0 * Y + Z, lambda X, Y, Z: -2.0 * X + 5.0 * Z, lambda X, Y, Z: 4.0 * X]
    U = [field_of(g, 1 << d, fns[d]) for d in range(3)]
    t = S.viscous_dissipation_term(g, S.cell_center_velocity(g, U))
    np.testing.assert_allclose(t[1:-1, 1:-1, 1:-1], 55.0, rtol=1e-9)
    # stretched grid: the collocated first derivative is still exact for linear fields
    g2 = grid(Ns=(7, 8, 9))
    U2 = [field_of(g2, 1 << d, fns[d]) for d in range(3)]
    t2 = S.viscous_dissipation_term(g2, S.cell_center_velocity(g2, U2))
    np.testing.assert_allclose(deep(t2), 55.0, rtol=1e-9)


def test_joule_heating_of_linear_edge_currents():
    # J_x = y, J_y = z, J_z = x on their edges: at the cell centres (y_c, z_c, x_c),
    # times sigmaInv_CC = 1 + x_c: (y_c^2 + z_c^2 + x_c^2)(1 + x_c)
    g = grid()
    fns = [lambda X, Y, Z: Y, lambda X, Y, Z: Z, lambda X, Y, Z: X]
    J = [field_of(g, S.EDGE_DL[d], fns[d]) for d in range(3)]
    sig = field_of(g, 0, lambda X, Y, Z: 1.0 + X)
    t = S.joule_heating_term(g, J, sig)
    ref = field_of(g, 0, lambda X, Y, Z: (Y * Y + Z * Z + X * X) * (1.0 + X))
    np.testing.assert_allclose(t, ref, rtol=1e-13, atol=1e-14)
    Jc = S.edge2cellcenter(g, J)
    for d in range(3):
        np.testing.assert_allclose(Jc[d], field_of(g, 0, fns[d]), rtol=0, atol=1e-14)


def test_volumetric_heating_profile_closed_form():
    g = grid()
    q = S.volumetric_heating_profile(g, 0.7)
    yc = g.coord(1, 1)
    F = 0.7 * 2.0 / (1.0 - np.exp(-2.0))
    for j, y in enumerate(yc):
        np.testing.assert_allclose(q[:, j, :], F * np.exp(-(y + 1.0)), rtol=1e-15)


def heated_state(g, method, rk_stages=1, diffusion=(False, 1.0), T0=None, qscale=0.7):
    T = g.zeros(0) if T0 is None else T0
    k0 = faces(g, lambda d, s: np.zeros(s))
    return S.EnergyState(g, T, O.BCs.make(["Neumann"] * 6), k0, dt=1e-2, method=method, advection=(False, 0.0),
                         diffusion=diffusion, isp=(50, 1e-14, 1e-300, 1), all_neumann=False,
                         volumetric_heating=(True, qscale), rk_stages=rk_stages)


def test_bdf2_grows_linearly_under_a_constant_source():
    # k = 0 (identity operator), constant source q: the AB2 start-up (Fnm1 = 0)
    # and BDF2's 4/3, -1/3 weights give T_n = n dt q exactly in closed form
    g = ugrid((6, 5, 4))
    st = heated_state(g, 4)
    q = S.volumetric_heating_profile(g, 0.7) * 0.7
    U = [g.zeros(1 << d) for d in range(3)]
    for n in range(1, 5):
        st.step(U)
        np.testing.assert_allclose(interior(st.T), interior(n * 1e-2 * q), rtol=1e-12)


def test_rk3_step_integrates_a_constant_source_exactly():
    # gamma + zeta sum to one over the three stages: one step adds dt q
    g = ugrid((6, 5, 4))
    st = heated_state(g, 5, rk_stages=3)
    q = S.volumetric_heating_profile(g, 0.7) * 0.7
    U = [g.zeros(1 << d) for d in range(3)]
    st.step(U)
    np.testing.assert_allclose(interior(st.T), interior(1e-2 * q), rtol=1e-12)
    st.step(U)
    np.testing.assert_allclose(interior(st.T), interior(2e-2 * q), rtol=1e-12)
    # inactive RK: all coefficients one, so F dt + Fnm1 dt: (1 + 0) then (1 + 1)
    s1 = heated_state(g, 5, rk_stages=1)
    s1.step(U)
    np.testing.assert_allclose(interior(s1.T), interior(1e-2 * q), rtol=1e-12)
    s1.step(U)
    np.testing.assert_allclose(interior(s1.T), interior(3e-2 * q), rtol=1e-12)


def test_rk3_explicit_diffusion_goes_through_L_with_alpha():
    # T = x^2 on a uniform grid, diffusion scale s, k = 0: lap T = 2 away from
    # the walls at every stage, so one step adds dt s 2 (alpha_1 + alpha_2 + alpha_3) = dt s
    g = ugrid((16, 5, 5))
    T0 = field_of(g, 0, lambda X, Y, Z: X * X)
    st = heated_state(g, 5, rk_stages=3, diffusion=(True, 0.3), T0=T0, qscale=0.0)
    U = [g.zeros(1 << d) for d in range(3)]
    st.step(U)
    inner = (slice(5, -5), slice(1, -1), slice(1, -1))
    np.testing.assert_allclose(st.T[inner] - T0[inner], 1e-2 * 0.3 * 2.0 * 0.5, rtol=1e-9)
    # L holds the last stage's diffusion
    np.testing.assert_allclose(st.L[inner], 0.3 * 2.0, rtol=1e-9)
