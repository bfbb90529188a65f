"""Known answers that pin the composition of the oracle's time step, each
derived from the mathematics rather than from the oracle's own formulas.

* Taylor-Green vortex (code/handwritten/grid_field/GF_distributions.f90:720-771,
  code/handwritten/solvers/Taylor_Green_Vortex_test.f90:32-98) on a periodic
  box [0, 2pi]^2 x [0, 1].  The velocity u = cos x sin y, v = -sin x cos y is
  an exact eigenvector of the discrete staggered Laplacian (eigenvalue
  lam = 2 (2 cos h - 2) / h^2) and is discretely divergence-free.  At a tiny
  amplitude the convection term (quadratic) is below round-off, so the
  amplitude follows a closed-form scalar recurrence:
    method 6 (backward Euler, coefficient dt nu alpha): A(n+1) = A(n) / (1 - dt nu lam).
  A wrong 1/Re or implicit coefficient moves A by at least 1e-6 relative.
* AB2 (method 4): a plane shear flow u = z (1 - z) between walls has zero
  convection and an exact collocated Laplacian on every row, so its source is
  the constant -2/Re: step 1 moves u by 1.5 dt F (Fnm1 = 0 at start,
  add_all_momentum_sources.f90:36-38), later steps by (1.5 - 0.5) dt F.
  Together they fix both AB2 weights and the start-up rule.
* At amplitude 1 the convection term of the vortex is (to O(h^2)) the
  gradient of its pressure -1/4 (cos 2x + cos 2y) F^2, so the projected
  pressure of method 6 (scale 1/dt, clean_divergence.f90:37-59) is AB2 of
  that pressure (1.5 of it at the first step); a wrong projection scale or
  start-up rule is off by a factor.
* J x B scales: J = curl(B) / Rem and J x B = N J x B0 with N = Ha^2 / Re
  (induction_sources.f90:70-84; the MHD flow configuration's
  SP%MT%JCrossB scale N and SP%IT%current scale 1/Rem).
* Implicit induction (method 6): with U = 0 the field B_x = sin z is an
  eigenvector of curl(curl) with eigenvalue mu = (2 - 2 cos h) / h^2, so one
  backward-Euler step divides it by 1 + dt mu / Rem
  (coeff_natural = -1/Rem, alpha = 1).
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import step_oracle as S

PER = ["Periodic"] * 6


def tgv_grid(N, NZ=4):
    two_pi = 2.0 * np.pi
    return O.Grid(O.uniform_nodes(0.0, two_pi, N), O.uniform_nodes(0.0, two_pi, N), O.uniform_nodes(0.0, 1.0, NZ))


def coords(g, dl):
    return np.meshgrid(*[g.coord(d, 0) if (dl >> d) & 1 else g.coord(d, 1) for d in range(3)], indexing="ij")


def tgv_mode(g, amp):
    X, Y, _ = coords(g, 1)
    u = np.asfortranarray(amp * np.cos(X) * np.sin(Y))
    X, Y, _ = coords(g, 2)
    v = np.asfortranarray(-amp * np.sin(X) * np.cos(Y))
    w = g.zeros(4)
    return [u, v, w]


def tgv_state(g, amp, alpha, Re, dt, isp_p=(200, 1e-13, 1e-300, 1)):
    U = tgv_mode(g, amp)
    bcs = [O.BCs.make(PER)] * 3
    return S.MHDState(g, U, bcs, O.BCs.make(PER), Re=Re, dt=dt, alpha_U=alpha,
                      isp_U=(200, 1e-14, 1e-300, 1), isp_p=isp_p)


def amplitude(g, st, amp0):
    mode = tgv_mode(g, 1.0)
    num = sum(float(np.sum(st.U[d][1:-1, 1:-1, 1:-1] * mode[d][1:-1, 1:-1, 1:-1])) for d in range(2))
    den = sum(float(np.sum(mode[d][1:-1, 1:-1, 1:-1] ** 2)) for d in range(2))
    return num / den


def residual_off_mode(g, st, A):
    mode = tgv_mode(g, A)
    return max(float(np.abs(st.U[d] - mode[d])[1:-1, 1:-1, 1:-1].max()) for d in range(3))


def test_tgv_mode_is_discrete_eigenvector():
    """The premise of the backward-Euler check: the staggered Laplacian of the
    implicit operator (lap_centered with periodic implicit BCs) maps the
    vortex to lam times it on every interior point, and its divergence is
    zero.  (The explicit collocated Laplacian does not: its wall rows use
    one-sided stencils on every BC type, stencils.f90:73-83.)"""
    N = 16
    g = tgv_grid(N)
    h = 2.0 * np.pi / N
    lam = 2.0 * (2.0 * np.cos(h) - 2.0) / (h * h)
    U = tgv_mode(g, 1.0)
    per = O.BCs.make(PER)
    for d in range(2):
        O.apply_bcs(g, U[d], 1 << d, per)
    Ax = O.op_vf(g, U, O.MOM_DIFFUSION, 1.0, [per] * 3, implicit=True)
    for d in range(2):
        # the operator is x + c lap(x); c = 1
        lapu = Ax[d] - U[d]
        np.testing.assert_allclose(lapu[1:-1, 1:-1, 1:-1][:-1 if d == 0 else None, :-1 if d == 1 else None],
                                   (lam * U[d])[1:-1, 1:-1, 1:-1][:-1 if d == 0 else None, :-1 if d == 1 else None],
                                   atol=1e-12)
    dv = O.div(g, U[0], U[1], U[2], 0)[1:-1, 1:-1, 1:-1]
    assert np.abs(dv).max() < 1e-14


def test_tgv_backward_euler_decay_closed_form():
    """Method 6: A(n+1) = A(n) / (1 - dt nu lam) exactly (tiny amplitude)."""
    N, Re, dt, A0 = 16, 50.0, 2e-2, 1e-12
    g = tgv_grid(N)
    h = 2.0 * np.pi / N
    lam = 2.0 * (2.0 * np.cos(h) - 2.0) / (h * h)
    nu = 1.0 / Re
    st = tgv_state(g, A0, 1.0, Re, dt)
    A = A0
    for n in range(6):
        st.step()
        A = A / (1.0 - dt * nu * lam)
        got = amplitude(g, st, A0)
        assert abs(got - A) <= 1e-11 * A0, (n, got, A, (got - A) / A0)
        # the velocity stays the pure mode (no other component appears)
        assert residual_off_mode(g, st, got) <= 1e-11 * A0
    assert (A0 - A) / A0 > 1e-3


def shear_state(Re, dt, alpha=0.0, N=6, NZ=16):
    """u = z (1 - z) between Dirichlet walls at z = 0, 1; x, y periodic.
    Its convection term is zero exactly (v = w = 0, u constant along x), and
    the collocated Laplacian of a quadratic is exact away from the wall rows
    (those use the wall value as a node one spacing away, stencils.f90:60-83),
    so F = (1/Re) u'' = -2/Re there.  The flow stays a function of z alone,
    so it stays divergence-free and the projection is the identity: a wall
    row's deviation reaches one more row per explicit step, no further."""
    g = O.Grid(O.uniform_nodes(0.0, 1.0, N), O.uniform_nodes(0.0, 1.0, N), O.uniform_nodes(0.0, 1.0, NZ))
    X, Y, Z = coords(g, 1)
    u = np.asfortranarray(Z * (1.0 - Z))
    wall = O.BCs.make(["Periodic"] * 4 + ["Dirichlet"] * 2)
    ubc = [wall, wall, wall]
    O.apply_bcs(g, u, 1, wall)
    st = S.MHDState(g, [u, g.zeros(2), g.zeros(4)], ubc, O.BCs.make(["Periodic"] * 4 + ["Neumann"] * 2),
                    Re=Re, dt=dt, alpha_U=alpha, isp_U=(200, 1e-14, 1e-300, 1), isp_p=(200, 1e-12, 1e-300, 1))
    return g, u, st


def test_ab2_closed_form_explicit_diffusion():
    """Method 4 under a constant source F = -2/Re: AB2 with Fnm1 = 0 at the
    start moves u by dt (1.5 F) in step 1 and by dt (1.5 - 0.5) F after."""
    Re, dt = 20.0, 1e-2
    g, u0, st = shear_state(Re, dt)
    F = -2.0 / Re
    for n in range(1, 5):
        st.step()
        shift = dt * 1.5 * F + (n - 1) * dt * 1.0 * F
        core = (slice(1, -1), slice(1, -1), slice(1 + n, -1 - n))  # rows the wall rows have not reached
        np.testing.assert_allclose(st.F[0][core], F, rtol=1e-10)
        np.testing.assert_allclose(st.U[0][core], (u0 + shift)[core], rtol=0, atol=1e-13)
        for d in (1, 2):
            assert np.abs(st.U[d][1:-1, 1:-1, 1:-1]).max() <= 1e-14


def test_planted_mistakes_fail_the_closed_forms():
    """Mistakes planted in the oracle's step (on its instance, on the CPU) are
    caught: a wrong implicit coefficient breaks the vortex decay, a wrong 1/Re
    or a non-zero Fnm1 at start-up breaks the AB2 shift."""
    N, Re, dt, A0 = 16, 50.0, 2e-2, 1e-12
    g = tgv_grid(N)
    h = 2.0 * np.pi / N
    lam = 2.0 * (2.0 * np.cos(h) - 2.0) / (h * h)
    for wrong in (-(1.0 / Re) * 0.5, -Re, -1.0 / Re * dt):
        st = tgv_state(g, A0, 1.0, Re, dt)
        st.coeff_implicit_U = wrong
        st.step()
        expect = A0 / (1.0 - dt * lam / Re)
        assert abs(amplitude(g, st, A0) - expect) > 1e-6 * A0
    Re, dt = 20.0, 1e-2
    for plant in ("Re", "start"):
        g2, u0, st = shear_state(Re, dt)
        if plant == "Re":
            st.coeff_explicit_U = Re
        else:
            st.F = [np.asfortranarray(np.full(f.shape, -2.0 / Re)) for f in st.F]
        st.step()
        shift = dt * 1.5 * (-2.0 / Re)
        assert np.abs(st.U[0] - (u0 + shift))[1:-1, 1:-1, 2:-2].max() > 1e-6


def test_tgv_pressure_follows_projection_scale():
    """Amplitude 1, method 6.  The vortex's convection term is (to O(h^2)) the
    gradient of its pressure P(t) = -1/4 (cos 2x + cos 2y) exp(-4 t / Re), so
    the projected pressure of step n (scale 1/dt) is AB2 of it: 1.5 P(0) at
    the first step (Fnm1 = 0), 1.5 P(t_n) - 0.5 P(t_n-1) after."""
    N, Re, dt = 32, 100.0, 1e-3
    g = tgv_grid(N)
    st = tgv_state(g, 1.0, 1.0, Re, dt, isp_p=(2000, 1e-10, 1e-300, 1))
    X, Y, _ = coords(g, 0)

    def P(t):
        return (-0.25 * (np.cos(2 * X) + np.cos(2 * Y)) * np.exp(-4.0 * t / Re))[1:-1, 1:-1, 1:-1]

    for n in range(3):
        st.step()
        pe = 1.5 * P(0.0) if n == 0 else 1.5 * P(n * dt) - 0.5 * P((n - 1) * dt)
        p = st.p[1:-1, 1:-1, 1:-1]
        p = p - p.mean()  # the periodic PPE fixes p up to a constant
        err = np.linalg.norm(p - pe) / np.linalg.norm(pe)
        assert err < 1e-2, (n, err)
    # the amplitude of the velocity decays like the continuum exp(-2 t / Re) to O(h^2)
    A = amplitude(g, st, 1.0)
    assert abs(A - np.exp(-2.0 * 3 * dt / Re)) < 1e-4


def test_jcrossb_scales_with_ha2_over_re_and_one_over_rem():
    """B = (0, 0, y): curl B = (1, 0, 0).  With Rem = 0.5 J_x = 2; with
    Ha = 3, Re = 2 (N = 4.5), J x B0 = (0, 0, N J_x) = (0, 0, 9).  The
    momentum source F then carries that J x B unscaled (U = 0)."""
    N = 8
    g = O.Grid(*[O.uniform_nodes(0.0, 1.0, N)] * 3)
    X, Y, Z = coords(g, 4)
    B = [g.zeros(1), g.zeros(2), np.asfortranarray(Y.copy())]
    Ubcs = [O.BCs.make(["Dirichlet"] * 6)] * 3
    st = S.MHDState(g, [g.zeros(1 << d) for d in range(3)], Ubcs, O.BCs.make(["Neumann"] * 6), Re=2.0, dt=1e-3,
                    B=B, Bbcs=[O.BCs.make(["Neumann"] * 6)] * 3, phibcs=O.BCs.make(["Dirichlet"] * 6),
                    Rem=0.5, Ha=3.0)
    np.testing.assert_allclose(st.J[0][2:-2, 2:-2, 2:-2], 2.0, rtol=1e-12)
    np.testing.assert_allclose(st.jCrossB[2][2:-2, 2:-2, 2:-2], 9.0, rtol=1e-12)
    np.testing.assert_allclose(st.jCrossB[0][2:-2, 2:-2, 2:-2], 0.0, atol=1e-12)
    st.momentum_sources()
    np.testing.assert_allclose(st.F[2][2:-2, 2:-2, 2:-2], 9.0, rtol=1e-12)


@pytest.mark.parametrize("Rem", [1.0, 0.25])
def test_induction_backward_euler_decay(Rem):
    """U = 0, Ha = 0, B_x = eps sin z on a periodic box: after n steps
    B_x = eps sin z / (1 + dt mu / Rem)^n; phi cleaning changes nothing (its
    right-hand side is round-off, under tol_abs)."""
    N, NZ, dt, eps = 6, 16, 5e-3, 1e-3
    g = O.Grid(O.uniform_nodes(0.0, 1.0, N), O.uniform_nodes(0.0, 1.0, N), O.uniform_nodes(0.0, 2 * np.pi, NZ))
    h = 2 * np.pi / NZ
    mu = (2.0 - 2.0 * np.cos(h)) / (h * h)
    X, Y, Z = coords(g, 1)
    Bx = np.asfortranarray(eps * np.sin(Z))
    per = O.BCs.make(PER)
    Ubcs = [O.BCs.make(["Dirichlet"] * 6)] * 3
    st = S.MHDState(g, [g.zeros(1 << d) for d in range(3)], Ubcs, O.BCs.make(["Neumann"] * 6), Re=100.0, dt=dt,
                    B=[Bx, g.zeros(2), g.zeros(4)], Bbcs=[per] * 3, phibcs=per, Rem=Rem, Ha=0.0,
                    isp_B=(200, 1e-14, 1e-300, 1), isp_phi=(200, 1e-10, 1e-15, 1))
    for n in range(1, 4):
        st.step()
        expect = Bx / (1.0 + dt * mu / Rem) ** n
        np.testing.assert_allclose(st.B[0][1:-1, 1:-1, 1:-1], expect[1:-1, 1:-1, 1:-1], rtol=0, atol=1e-13 * eps)
        for d in (1, 2):
            assert np.abs(st.B[d][1:-1, 1:-1, 1:-1]).max() <= 1e-15
    # a mistake planted in the oracle's coefficient (Rem for 1/Rem, or 1/Rem^2) is caught
    for wrong in (Rem + 1.0, 1.0 / Rem ** 2 + 1.0):
        st.coeff_implicit_B = wrong
        st.step()
        expect = expect / (1.0 + dt * mu / Rem)
        assert np.abs(st.B[0] - expect)[1:-1, 1:-1, 1:-1].max() > 1e-8 * eps


def mpg_state(Re, dt, scale, N=6, NZ=12):
    """Fluid at rest between walls at z = 0, 1 (x, y periodic), driven along x
    by the mean pressure gradient, method 4."""
    g = O.Grid(O.uniform_nodes(0.0, 1.0, N), O.uniform_nodes(0.0, 1.0, N), O.uniform_nodes(0.0, 1.0, NZ))
    wall = O.BCs.make(["Periodic"] * 4 + ["Dirichlet"] * 2)
    st = S.MHDState(g, [g.zeros(1 << d) for d in range(3)], [wall] * 3,
                    O.BCs.make(["Periodic"] * 4 + ["Neumann"] * 2), Re=Re, dt=dt, alpha_U=0.0,
                    isp_p=(200, 1e-12, 1e-300, 1), mpg_dir=1, mpg_scale=scale)
    return g, st


def test_mean_pressure_gradient_forcing():
    """compute_add_MPG adds scale dt to u before the sources are computed:
    away from the walls the flow is uniform along x (no convection, zero
    Laplacian, divergence-free), so after n steps u = n scale dt exactly on the
    rows the walls have not reached; on the wall rows the first step's source
    is already the diffusion of the forced u (so the forcing precedes the
    sources).  Planted mistakes (twice the scale, the forcing after the
    sources) fail."""
    Re, dt, s = 10.0, 1e-2, 3.0
    g, st = mpg_state(Re, dt, s)
    for n in range(1, 4):
        st.step()
        core = (slice(1, -1), slice(1, -1), slice(1 + n, -1 - n))
        np.testing.assert_allclose(st.U[0][core], n * s * dt, rtol=1e-13)
        if n == 1:
            assert np.abs(st.F[0][1:-1, 1:-1, 1]).min() > 1e-3 * s * dt
            assert np.abs(st.F[0][1:-1, 1:-1, 3:-3]).max() <= 1e-12
        for d in (1, 2):
            assert np.abs(st.U[d][1:-1, 1:-1, 1:-1]).max() <= 1e-14
    g, st = mpg_state(Re, dt, s)
    st.mpg_scale = 2 * s
    st.step()
    assert np.abs(st.U[0][1:-1, 1:-1, 3:-3] - s * dt).max() > 1e-3 * s * dt
    g, st = mpg_state(Re, dt, s)
    st.mpg_dir = 0  # forcing applied after the sources instead
    st.momentum_sources()
    st.U[0] = np.asfortranarray(st.U[0] + s * dt)
    assert np.abs(st.F[0][1:-1, 1:-1, 1]).max() == 0.0


def test_finite_rem_btot_enters_jcrossb():
    """finite_Rem: Btot = B0 + B (compute_Btot).  B = (0, 0, y), B0 = (0, 1, 0):
    J = (1/Rem, 0, 0) and J x Btot = N (0, -y/Rem, 1/Rem); with Btot = B0 the
    y component vanishes."""
    N = 8
    g = O.Grid(*[O.uniform_nodes(0.0, 1.0, N)] * 3)
    X, Y, Z = coords(g, 4)
    Ubcs = [O.BCs.make(["Dirichlet"] * 6)] * 3
    Rem, Ha, Re = 0.5, 3.0, 2.0
    Nf = Ha * Ha / Re
    for finite in (True, False):
        B = [g.zeros(1), g.zeros(2), np.asfortranarray(Y.copy())]
        st = S.MHDState(g, [g.zeros(1 << d) for d in range(3)], Ubcs, O.BCs.make(["Neumann"] * 6), Re=Re, dt=1e-3,
                        B=B, Bbcs=[O.BCs.make(["Neumann"] * 6)] * 3, phibcs=O.BCs.make(["Dirichlet"] * 6),
                        Rem=Rem, Ha=Ha, finite_Rem=finite)
        _, Yy, _ = coords(g, 2)
        jy = st.jCrossB[1][2:-2, 2:-2, 2:-2]
        expect = -Nf * Yy[2:-2, 2:-2, 2:-2] / Rem if finite else 0.0 * Yy[2:-2, 2:-2, 2:-2]
        np.testing.assert_allclose(jy, expect, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(st.jCrossB[2][2:-2, 2:-2, 2:-2], Nf / Rem, rtol=1e-12)


def test_robin_rules_satisfy_u_plus_cw_dudx():
    """The thin-wall rules impose u + c_w du/dx = 0 at every face, with du/dx
    along the coordinate (not the outward normal): cell-centred data through
    the wall value (ghost + interior)/2 and the one-sided difference over dh
    (the value term is dropped, as in the reference), node-centred data at the
    boundary node with the central difference over 2 dh, where the value b
    gives u + c_w du/dx = b."""
    rng = np.random.default_rng(11)
    hs = [O.roberts_both_nodes(0, 1, 7, 1.1), O.uniform_nodes(-1, 1, 6), O.roberts_both_nodes(0, 2, 9, 1.05)]
    g = O.Grid(*hs)
    cw = [0.01, 0.3, 2.0, 0.05, 0.7, 0.02]
    for dl in (0, 1, 2, 4, 7):
        for implicit in (False, True):
            vals = [0.4, -0.3, 0.2, 0.1, -0.6, 0.5]
            u = np.asfortranarray(rng.standard_normal(g.shape(dl)))
            O.apply_bcs(g, u, dl, O.BCs.make(["Robin"] * 6, vals, cw), implicit)
            for face in range(6):
                d, is_max = face // 2, face % 2
                um = np.moveaxis(u, d, 0)
                s = um.shape[0]
                dh = O.face_dh(g, d, is_max)
                core = (slice(2, -2), slice(2, -2))  # away from the edges other faces wrote
                if not (dl >> d) & 1:
                    G, I = (s - 1, s - 2) if is_max else (0, 1)
                    uw = (um[G] + um[I]) / 2
                    dudx = (um[G] - um[I]) / dh if is_max else (um[I] - um[G]) / dh
                    target = 0.0
                else:
                    G, B, I = (s - 1, s - 2, s - 3) if is_max else (0, 1, 2)
                    uw = um[B]
                    dudx = (um[G] - um[I]) / (2 * dh) if is_max else (um[I] - um[G]) / (2 * dh)
                    target = 0.0 if implicit else vals[face]
                res = (uw + cw[face] * dudx - target)[core]
                scale = max(1.0, np.abs(cw[face] * dudx[core]).max())
                assert np.abs(res).max() <= 1e-13 * scale, (dl, face, implicit)
                # the outward-normal reading u + c_w du/dn = 0 fails at the min faces
                if not is_max:
                    assert np.abs((uw - cw[face] * dudx - target)[core]).max() > 1e-6


def test_buoyancy_and_gravity_closed_form():
    """compute_add_buoyancy / compute_add_gravity (momentum_sources.f90:133-167)
    with U = 0 (no other source), uniform gravity along y (the reference's
    uniform_gravity_field, 1 in g_y everywhere) and T linear in y on a uniform
    grid: cellCenter2Face is exact on linear data, so on every y-face but the
    ghost planes F_y = Gr/Re^2 * T(y_face) + 1/Fr^2, F_x = F_z = 0 there, and the
    ghost planes are 0 (extractFace)."""
    N = (6, 8, 5)
    hs = [O.uniform_nodes(0.0, 1.0, N[0]), O.uniform_nodes(-1.0, 1.0, N[1]), O.uniform_nodes(0.0, 2.0, N[2])]
    g = O.Grid(*hs)
    U0 = [g.zeros(1 << d) for d in range(3)]
    bcs = [O.BCs.make(["Dirichlet"] * 6) for _ in range(3)]
    yc = g.coord(1, 1)
    T = np.asfortranarray(np.broadcast_to((0.75 + 2.5 * yc)[None, :, None], g.shape(0)).copy())
    grav = [g.zeros(0), np.asfortranarray(np.ones(g.shape(0))), g.zeros(0)]
    sb, sg = 0.37, -1.9
    st = S.MHDState(g, U0, bcs, O.BCs.make(["Neumann"] * 6), Re=100.0, dt=1e-3, alpha_U=1.0, T=T, gravity=grav,
                    buoyancy_scale=sb, gravity_scale=sg)
    st.momentum_sources()
    Fy = st.F[1]
    yf = g.coord(1, 0)
    expect = sb * (0.75 + 2.5 * yf) + sg
    inner = Fy[1:-1, 1:-1, 1:-1]
    np.testing.assert_allclose(inner, np.broadcast_to(expect[None, 1:-1, None], inner.shape), rtol=0, atol=1e-13)
    for F in (Fy, st.F[0], st.F[2]):
        for ax in range(3):
            for idx in (0, -1):
                assert not np.take(F, idx, axis=ax).any()
    assert not st.F[0].any() and not st.F[2].any()
    # buoyancy alone: the gravity term off leaves the T term only
    st2 = S.MHDState(g, U0, bcs, O.BCs.make(["Neumann"] * 6), Re=100.0, dt=1e-3, alpha_U=1.0, T=T, gravity=grav,
                     buoyancy_scale=sb)
    st2.momentum_sources()
    np.testing.assert_allclose(st2.F[1][1:-1, 1:-1, 1:-1],
                               np.broadcast_to((sb * (0.75 + 2.5 * yf))[None, 1:-1, None], inner.shape), atol=1e-13)
