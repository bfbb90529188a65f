"""FP64 restatement of one MOONS MHD time step (momentum + low-Rem induction),
composed from the C oracle's operators.

TEST INFRASTRUCTURE ONLY: the checker for the device time step.

Sequence (code/handwritten/MOONS/MOONS_solver.f90:47-144, one RK stage):
  momentum:  add_all_momentum_sources (solvers/momentum/add_all_momentum_sources.f90:22-86)
               Fnm1 = F; L = 0; F = 0
               F += -1 * u.grad(u)        advect_U_convection1 (ops/ops_advect.f90:71-127)
               F += coeff_explicit lap(U) compute_add_diffusion (momentum_sources.f90:87-95),
                                          lapUniformCoeff_VF (ops/ops_discrete.f90:265-273)
               F += 1 * jCrossB           extract_add_JCrossB (momentum_sources.f90:111-121)
               F += buoyancy * scale      compute_add_buoyancy (momentum_sources.f90:133-150):
                                          cellCenter2Face(T g), extractFace
               F += gravity * scale       compute_add_gravity (momentum_sources.f90:152-167)
             solve_momentum (governing_equations/momentum_extend.f90:326-374)
               method 4: Euler_time_no_diff_AB2_sources (solvers/time_marching_methods.f90:266-296)
               method 6: Euler_time_AB2_sources (time_marching_methods.f90:117-150)
               U_E = face2Edge_no_diag(U)
  induction: add_all_induction_sources (solvers/induction/add_all_induction_sources.f90:24-60)
               Fnm1 = F; L = 0; F = 0
               F += 1 * curl(U_E x Btot)  add_curl_U_cross_B (induction_sources.f90:28-38)
               F += -beta * curl(sigInv J) add_curl_J (induction_sources.f90:40-49)
             solve_induction, method 6 (induction_extend.f90:475-525)
             J = curl(B)/Rem; Btot = B0 (finite_Rem = F); jCrossB = N edge2Face(J) x face2Face(Btot)
                                          (induction_sources.f90:70-84)
With one mesh, embedding and extraction between the fluid and induction
meshes (embedVelocity_E, extractFace) are the identity.  With an induction
grid gi and the fluid's Domain in it (oracle/domain_oracle.py), B, phi, J and
sigma_inv live on gi, U_E is embedded into gi each step (zero outside the
fluid; add_all_induction_sources.f90:35-36) and J x B is extracted to the fluid
faces (extract_add_JCrossB, momentum_sources.f90:111-121).
Element-wise arithmetic follows the reference's order (multiply then add,
AB2 = 1.5 Fn then + (-0.5) Fnm1).
"""
from __future__ import annotations

import numpy as np

from . import domain_oracle as D
from . import oracle as O

EDGE_DL = (6, 5, 3)
# (component, location) of the six off-diagonal entries, in the C ABI's order:
# U_y@x-edges, U_z@x, U_x@y, U_z@y, U_x@z, U_y@z
NO_DIAG = ((1, 0), (2, 0), (0, 1), (2, 1), (0, 2), (1, 2))


def face2edge_no_diag(g, U):
    """face2Edge_no_diag_TF (ops/ops_interp.f90:531-544): U_i onto j-edges, one
    interpolation along the third direction."""
    return {(i, j): O.interp(g, U[i], 1 << i, EDGE_DL[j], 3 - i - j) for (i, j) in NO_DIAG}


def face2face(g, u, i, j):
    """face2Face_SF (ops_interp.f90:287-295): face i -> CC -> face j."""
    cc = O.interp(g, u, 1 << i, 0, i)
    return O.interp(g, cc, 0, 1 << j, j)


def edge2face(g, e, i, j):
    """edge2Face_SF (ops_interp.f90:400-413): edge i -> face j along the third direction."""
    return O.interp(g, e, EDGE_DL[i], 1 << j, 3 - i - j)


def advect_U_convection(g, U, U_E):
    """div_i = u_j d/dx_j u_i (ops/ops_advect.f90:71-127)."""
    div = []
    for d in range(3):
        cc = O.interp(g, U[d], 1 << d, 0, d)
        t = g.zeros(1 << d)
        O.del_op(g, t, 1 << d, cc, 0, 1, d, 0, 1)
        div.append(t * U[d])
    # (edge entry, derivative direction, velocity component) per component, reference order
    terms = {0: (((0, 2), 1, 1), ((0, 1), 2, 2)),
             1: (((1, 2), 0, 0), ((1, 0), 2, 2)),
             2: (((2, 1), 0, 0), ((2, 0), 1, 1))}
    for d in range(3):
        for key, ddir, vel in terms[d]:
            t1 = g.zeros(1 << d)
            O.del_op(g, t1, 1 << d, U_E[key], EDGE_DL[key[1]], 1, ddir, 0, 1)
            t2 = face2face(g, U[vel], vel, d)
            t1 = t1 * t2
            div[d] = div[d] + t1
    return div


def lap_collocated(g, U):
    """lapUniformCoeff_VF: second derivatives along x, y, z (assign, add, add)."""
    out = []
    for d in range(3):
        t = g.zeros(1 << d)
        O.del_op(g, t, 1 << d, U[d], 1 << d, 2, 0, 0, 1)
        O.del_op(g, t, 1 << d, U[d], 1 << d, 2, 1, 0, 2)
        O.del_op(g, t, 1 << d, U[d], 1 << d, 2, 2, 0, 2)
        out.append(t)
    return out


def zero_ghosts(g, u, dl):
    """assign_ghost_XPeriodic(u, 0) for a field with no BCs: all six ghost planes."""
    u = u.copy(order="F")
    lib = O.lib()
    lib.orc_assign_ghost_xperiodic(O._F(u), dl, g.h, None, 0.0)
    return u


def assign_wall_dirichlet(g, u, dl, bcs):
    u = u.copy(order="F")
    O.lib().orc_assign_wall_dirichlet(O._F(u), dl, g.h, O.C.byref(bcs), 0.0)
    return u


def clean_div(g, X, Xstar, xbcs, phi, phibcs, all_neumann, scale, isp):
    """clean_div_PCG (solvers/clean_divergence.f90:37-59)."""
    t = O.div(g, Xstar[0], Xstar[1], Xstar[2], 0)
    t = np.asfortranarray(t * scale)
    it, res0, res, _ = O.pcg_lap(g, 0, phibcs, all_neumann, phi, t, *isp)
    gr = O.grad(g, phi, 0)
    inv = 1.0 / scale
    for d in range(3):
        X[d] = np.asfortranarray(Xstar[d] - gr[d] * inv)
        O.apply_bcs(g, X[d], 1 << d, xbcs[d])
    return it


class MHDState:
    """State and parameters of the coupled step; arrays are F-ordered FP64."""

    finite_Rem = False
    mpg_dir, mpg_scale = 0, 1.0

    def __init__(self, g, U, Ubcs, pbcs, Re, dt, alpha_U=1.0, B=None, Bbcs=None, phibcs=None, Rem=1.0, Ha=0.0,
                 B0=(0.0, 1.0, 0.0), isp_U=(5, 1e-6, 1e-13, 1), isp_p=(5, 1e-6, 1e-13, 1),
                 isp_B=(5, 1e-6, 1e-13, 1), isp_phi=(5, 1e-6, 1e-13, 1), gi=None, dom=None, sigInv=None,
                 mpg_dir=0, mpg_scale=1.0, finite_Rem=False, T=None, gravity=None, buoyancy_scale=None,
                 gravity_scale=None):
        self.g = g
        # buoyancy (Gr/Re^2) and gravity (1/Fr^2) sources: the energy's T (CC) and
        # the CC gravity field (SP%MT%Buoyancy, SP%MT%Gravity; None = off)
        self.T = T
        self.gravity = gravity
        self.buoyancy_scale, self.gravity_scale = buoyancy_scale, gravity_scale
        # mean pressure gradient forcing (SP%MT%mean_pressure_grad, SP%SCP%mpg_dir; 0 = off)
        self.mpg_dir, self.mpg_scale = mpg_dir, mpg_scale
        self.finite_Rem = finite_Rem  # Btot = B0 + B (SP%SCP%finite_Rem), else B0
        self.gi = gi if gi is not None else g  # the induction grid
        self.dom = dom
        self.Re, self.dt, self.alpha_U = Re, dt, alpha_U
        self.coeff_explicit_U = (1.0 / Re) * (1.0 - alpha_U)
        self.coeff_implicit_U = -(1.0 / Re) * alpha_U
        self.U = [np.asfortranarray(u) for u in U]
        self.Ubcs, self.pbcs = Ubcs, pbcs
        self.p = g.zeros(0)
        self.F = [g.zeros(1 << d) for d in range(3)]
        self.Fnm1 = [g.zeros(1 << d) for d in range(3)]
        self.L = [g.zeros(1 << d) for d in range(3)]
        self.isp_U, self.isp_p, self.isp_B, self.isp_phi = isp_U, isp_p, isp_B, isp_phi
        self.induction = B is not None
        self.U_E = face2edge_no_diag(g, self.U)
        self.jCrossB = [g.zeros(1 << d) for d in range(3)]
        self.iters = {}
        if self.induction:
            self.Rem, self.N = Rem, Ha * Ha / Re
            self.coeff_implicit_B = 1.0 / Rem  # -coeff_natural * alpha, coeff_natural = -1/Rem, alpha = 1
            gi = self.gi
            self.B = [np.asfortranarray(b) for b in B]
            self.Bbcs, self.phibcs = Bbcs, phibcs
            self.phi = gi.zeros(0)
            self.B0 = [np.asfortranarray(np.full(gi.shape(1 << d), B0[d])) for d in range(3)]
            self.FB = [gi.zeros(1 << d) for d in range(3)]
            self.FBnm1 = [gi.zeros(1 << d) for d in range(3)]
            self.LB = [gi.zeros(1 << d) for d in range(3)]
            self.sigInv = ([np.asfortranarray(s) for s in sigInv] if sigInv is not None
                           else [np.asfortranarray(np.ones(gi.shape(dl))) for dl in EDGE_DL])
            self.U_Ei = {k: gi.zeros(EDGE_DL[k[1]]) for k in NO_DIAG}  # embedded U_E, zero outside the fluid
            self.post_induction()

    # ---------------------------------------------------------------- momentum
    def momentum_sources(self):
        g = self.g
        if self.mpg_dir:
            # compute_add_MPG (hw/solvers/momentum/momentum_sources.f90:97-108), first in
            # add_all_momentum_sources (add_all_momentum_sources.f90:31-33): every value of
            # the component, ghosts included, += scale*dt; then apply_BCs(U)
            d = self.mpg_dir - 1
            self.U[d] = np.asfortranarray(self.U[d] + self.mpg_scale * self.dt)
            for e in range(3):
                O.apply_bcs(g, self.U[e], 1 << e, self.Ubcs[e])
        self.Fnm1 = [f.copy(order="F") for f in self.F]
        self.L = [g.zeros(1 << d) for d in range(3)]
        F = [g.zeros(1 << d) for d in range(3)]
        adv = advect_U_convection(g, self.U, self.U_E)
        F = [F[d] + adv[d] * (-1.0) for d in range(3)]
        lap = lap_collocated(g, self.U)
        F = [F[d] + lap[d] * self.coeff_explicit_U for d in range(3)]
        if self.induction:
            if self.dom is None:
                jcb = [zero_ghosts(g, self.jCrossB[d], 1 << d) for d in range(3)]
            else:
                jcb = [zero_ghosts(g, D.extract(g.zeros(1 << d), self.jCrossB[d], 1 << d, self.dom), 1 << d)
                       for d in range(3)]
            F = [F[d] + jcb[d] * 1.0 for d in range(3)]
        if self.buoyancy_scale is not None:
            # compute_add_buoyancy (momentum_sources.f90:133-150): temp_CC = T, times
            # gravity, cellCenter2Face (interp along each component's direction),
            # extractFace onto the fluid faces (energy on the fluid mesh), F += b*scale
            for d in range(3):
                tcc = np.asfortranarray(self.T * self.gravity[d])
                b = zero_ghosts(g, O.interp(g, tcc, 0, 1 << d, d), 1 << d)
                F[d] = F[d] + b * self.buoyancy_scale
        if self.gravity_scale is not None:
            # compute_add_gravity (momentum_sources.f90:152-167)
            for d in range(3):
                gf = zero_ghosts(g, O.interp(g, np.asfortranarray(self.gravity[d]), 0, 1 << d, d), 1 << d)
                F[d] = F[d] + gf * self.gravity_scale
        self.F = [np.asfortranarray(f) for f in F]

    def _ab2_rhs(self, X, Xbcs, F, Fnm1, L, g=None):
        g = self.g if g is None else g
        out = []
        for d in range(3):
            t = F[d] * 1.5
            t = t + Fnm1[d] * (-0.5)
            t = t + L[d]
            t = t * self.dt
            t = assign_wall_dirichlet(g, np.asfortranarray(t), 1 << d, Xbcs[d])
            t = t + X[d]
            out.append(np.asfortranarray(t))
        return out

    def solve_momentum(self):
        g = self.g
        if self.alpha_U == 0.0:  # method 4: Euler_time_no_diff_AB2_sources (no L)
            Ustar = []
            for d in range(3):
                t = self.F[d] * 1.5
                t = t + self.Fnm1[d] * (-0.5)
                t = t * self.dt
                t = assign_wall_dirichlet(g, np.asfortranarray(t), 1 << d, self.Ubcs[d])
                Ustar.append(np.asfortranarray(t + self.U[d]))
            scale = 1.0
        else:  # method 6: Euler_time_AB2_sources
            rhs = self._ab2_rhs(self.U, self.Ubcs, self.F, self.Fnm1, self.L)
            Ustar = [u.copy(order="F") for u in self.U]
            c = self.dt * 1.0 * self.coeff_implicit_U
            it, _, _, _ = O.pcg_vf(g, O.MOM_DIFFUSION, c, self.Ubcs, Ustar, rhs, *self.isp_U)
            self.iters["U"] = it
            scale = 1.0 / self.dt
        self.iters["p"] = clean_div(g, self.U, Ustar, self.Ubcs, self.p, self.pbcs, True, scale, self.isp_p)
        self.Ustar = Ustar
        self.U_E = face2edge_no_diag(g, self.U)

    # --------------------------------------------------------------- induction
    def induction_sources(self):
        g = self.gi
        self.FBnm1 = [f.copy(order="F") for f in self.FB]
        self.LB = [g.zeros(1 << d) for d in range(3)]
        F = [g.zeros(1 << d) for d in range(3)]
        if self.dom is None:
            U_E = [self.U_E[k] for k in NO_DIAG]
        else:
            for k in NO_DIAG:
                self.U_Ei[k] = D.embed(self.U_Ei[k], self.U_E[k], EDGE_DL[k[1]], self.dom)
            U_E = [self.U_Ei[k] for k in NO_DIAG]
        adv = O.advect_B(g, U_E, self.Btot)
        F = [F[d] + adv[d] * 1.0 for d in range(3)]
        E = [np.asfortranarray(self.J[d] * self.sigInv[d]) for d in range(3)]
        cj = [O.curl(g, E[0], E[1], E[2], 1 << d, d) for d in range(3)]
        beta_B = 0.0
        F = [F[d] + cj[d] * (-beta_B) for d in range(3)]
        self.FB = [np.asfortranarray(f) for f in F]

    def solve_induction(self):
        g = self.gi
        rhs = self._ab2_rhs(self.B, self.Bbcs, self.FB, self.FBnm1, self.LB, g)
        Bstar = [b.copy(order="F") for b in self.B]
        c = self.dt * 1.0 * self.coeff_implicit_B
        it, _, _, _ = O.pcg_vf(g, O.IND_DIFFUSION, c, self.Bbcs, Bstar, rhs, *self.isp_B, k=self.sigInv)
        self.iters["B"] = it
        self.iters["phi"] = clean_div(g, self.B, Bstar, self.Bbcs, self.phi, self.phibcs, False, 1.0 / self.dt,
                                      self.isp_phi)
        self.post_induction()

    def post_induction(self):
        g = self.gi
        self.J = [np.asfortranarray(O.curl(g, self.B[0], self.B[1], self.B[2], EDGE_DL[d], d) * (1.0 / self.Rem))
                  for d in range(3)]
        # compute_Btot (hw/governing_equations/induction_extend.f90:527-536)
        if self.finite_Rem:
            self.Btot = [np.asfortranarray(self.B0[d] + self.B[d]) for d in range(3)]
        else:
            self.Btot = [b.copy(order="F") for b in self.B0]
        A = {(i, j): edge2face(g, self.J[i], i, j) for (i, j) in NO_DIAG}
        Bf = {(i, j): face2face(g, self.Btot[i], i, j) for (i, j) in NO_DIAG}
        jx = A[(1, 0)] * Bf[(2, 0)] - A[(2, 0)] * Bf[(1, 0)]
        jy = -(A[(0, 1)] * Bf[(2, 1)] - A[(2, 1)] * Bf[(0, 1)])
        jz = A[(0, 2)] * Bf[(1, 2)] - A[(1, 2)] * Bf[(0, 2)]
        self.jCrossB = [np.asfortranarray(j * self.N) for j in (jx, jy, jz)]

    def step(self):
        self.momentum_sources()
        self.solve_momentum()
        if self.induction:
            self.induction_sources()
            self.solve_induction()


# the 3-stage low-storage RK coefficients (init_RKP, code/handwritten/var_set/RK_params_extend.f90:29-62):
# gamma, zeta, alpha per stage, beta = alpha; all ones when RK is inactive
RK3 = {"gamma": (8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0), "zeta": (0.0, -17.0 / 60.0, -5.0 / 12.0),
       "alpha": (4.0 / 15.0, 1.0 / 15.0, 1.0 / 6.0)}
TWO_THIRDS, FOUR_THIRDS, NEG_ONE_THIRD = 2.0 / 3.0, 4.0 / 3.0, -1.0 / 3.0  # time_marching_methods_SF.f90:35-37


def cell_center_velocity(g, U):
    """face2CellCenter of each component (ops_interp.f90:232-240): N -> CC along d."""
    return [np.asfortranarray(O.interp(g, U[d], 1 << d, 0, d)) for d in range(3)]


def edge2cellcenter(g, J):
    """edge2CellCenter_VF (ops_interp.f90:688-696, edge2CellCenter_SF :415-431):
    component d goes edge -> face (faceDir = d+1 mod 3, one interpolation
    along the third direction) -> cell centre (along faceDir)."""
    out = []
    for d in range(3):
        fd = (d + 1) % 3
        od = 3 - d - fd
        face = O.interp(g, J[d], EDGE_DL[d], 1 << fd, od)
        out.append(np.asfortranarray(O.interp(g, face, 1 << fd, 0, fd)))
    return out


def lap_uniform_cc(g, f):
    """lapUniformCoeff_SF (ops_discrete.f90:101-110) of a cell-centred field:
    collocated second derivatives along x (assign), y and z (add)."""
    t = g.zeros(0)
    O.del_op(g, t, 0, f, 0, 2, 0, 0, 1)
    O.del_op(g, t, 0, f, 0, 2, 1, 0, 2)
    O.del_op(g, t, 0, f, 0, 2, 2, 0, 2)
    return t


def ke_diffusion_term(g, U_CC):
    """add_KE_diffusion (energy_sources.f90:50-63) before the scaling:
    lap(Ux^2 + Uy^2 + Uz^2); square(VF) then add(SF, VF) = x + y + z."""
    sq = [u * u for u in U_CC]
    t = np.asfortranarray(sq[0] + sq[1] + sq[2])
    return lap_uniform_cc(g, t)


def viscous_dissipation_term(g, U_CC):
    """add_viscous_dissipation (energy_sources.f90:65-76) before the scaling:
    grad_TF (ops_discrete.f90:313-321): G[c][e] = d U_c / d x_e, collocated
    (del n = 1, pad 0) on the cell centres; square(TF); add(SF, TF)
    (TF_extend.f90:248-255) sums G[x][x], G[y][x], G[z][x], G[x][y], ...,
    G[z][z] left to right (GF_add.f90:137-171)."""
    G = [[None] * 3 for _ in range(3)]
    for c in range(3):
        for e in range(3):
            t = g.zeros(0)
            O.del_op(g, t, 0, U_CC[c], 0, 1, e, 0, 1)
            G[c][e] = t
    acc = None
    for e in range(3):
        for c in range(3):
            sq = G[c][e] * G[c][e]
            acc = sq if acc is None else acc + sq
    return np.asfortranarray(acc)


def joule_heating_term(g, J, sigma_inv_cc):
    """add_joule_heating (energy_sources.f90:78-93) before the scaling:
    (Jx_cc^2 + Jy_cc^2 + Jz_cc^2) * sigmaInv_CC."""
    jc = edge2cellcenter(g, J)
    sq = [j * j for j in jc]
    t = sq[0] + sq[1] + sq[2]
    return np.asfortranarray(t * sigma_inv_cc)


def volumetric_heating_profile(g, scale):
    """add_volumetric_heating_Nuclear (energy_sources.f90:95-113) before its
    final scaling: F exp(-m (y_c + 1)) with a = L = 1, m = a/L,
    F = scale 2 m / (1 - exp(-2 m)) (scalar libm exp, as the library's host code)."""
    import math

    a, L = 1.0, 1.0
    mm = a / L
    Fc = scale * 2.0 * mm / (1.0 - math.exp(-2.0 * mm))
    hc = g.coord(1, 1)
    prof = np.array([Fc * math.exp(-mm * (hc[j] + 1.0)) for j in range(len(hc))])
    s = g.shape(0)
    return np.asfortranarray(np.broadcast_to(prof[None, :, None], s))


class EnergyState:
    """The energy equation of one step (solve_energy,
    governing_equations/energy_extend.f90:236-262, after
    add_all_energy_sources, solvers/energy/add_all_energy_sources.f90:21-80):

      Fnm1 = F; (RK active: L = 0); F = 0
      advection: F += scale * (-div(cellCenter2Face(T) * U))   add_advection (energy_sources.f90:26-38)
      diffusion: (RK active: L, else F) += scale * lap(T)      add_diffusion (energy_sources.f90:40-48)
      KE diffusion: F += scale * lap(|U_CC|^2)                 add_KE_diffusion (:50-63)
      viscous dissipation: F += scale * grad(U_CC):grad(U_CC)  add_viscous_dissipation (:65-76)
      Joule heating: F += scale * |J_CC|^2 sigmaInv_CC         add_joule_heating (:78-93)
      volumetric heating: F += scale * q(y)                    add_volumetric_heating_Nuclear (:95-113)
      method 1: T = T + F dt             Euler_time_no_diff_Euler_sources_SF (time_marching_methods_SF.f90:154-170)
      method 2: T = T + dt AB2(F, Fnm1)  Euler_time_no_diff_AB2_sources_SF (:135-152)
      method 3: Tnm1 = T; solve (I + c div(k grad)) T = T + dt AB2(F, Fnm1), c = dt coeff_implicit
                                         Euler_time_AB2_sources_SF (:71-98)
      method 4: rhs = 2/3 dt AB2(F, Fnm1) + 4/3 T - 1/3 Tnm1; Tnm1 = T; c = dt 2/3 coeff_implicit
                                         O2_BDF_time_AB2_sources_SF (:42-69)
      method 5: rhs = F dt gamma + Fnm1 dt zeta + L dt alpha + T; Tnm1 = T; c = dt beta coeff_implicit
                                         Euler_time_RK_sources_SF (:100-129)

    With rk_stages = 3 one step runs the three RK stages (sources, then the
    march with that stage's coefficients), as the stage loop of solve_MOONS
    (MOONS/MOONS_solver.f90:70-90); with 1 the coefficients are all one.
    U is the face velocity the step reads (embedFace and embedCC are the
    identity on one domain; U_CC = face2CellCenter(U), which the momentum
    solve leaves behind, momentum_extend.f90:370).  k are the face arrays bound
    to the T solver's operator.  The Dirichlet wall planes of
    assign_wall_Dirichlet do not exist on a cell-centred field, so that call
    changes nothing here.
    """

    def __init__(self, g, T, Tbcs, k, dt, method=3, coeff_implicit=-1.0, advection=(True, -1.0),
                 diffusion=(False, 1.0), isp=(5, 1e-6, 1e-13, 1), all_neumann=False, KE_diffusion=(False, 1.0),
                 viscous_dissipation=(False, 1.0), joule_heating=(False, 1.0), volumetric_heating=(False, 1.0),
                 rk_stages=1, J=None, sigma_inv_cc=None):
        self.g = g
        self.T = np.array(T, dtype=np.float64, order="F")  # a copy: the step advances it in place
        self.Tbcs, self.k = Tbcs, [np.asfortranarray(a) for a in k]
        self.dt, self.method, self.coeff_implicit = dt, method, coeff_implicit
        self.advection, self.diffusion, self.isp, self.all_neumann = advection, diffusion, isp, all_neumann
        self.KE_diffusion, self.viscous_dissipation = KE_diffusion, viscous_dissipation
        self.joule_heating, self.volumetric_heating = joule_heating, volumetric_heating
        self.rk_stages, self.stage = rk_stages, 0
        self.J, self.sigma_inv_cc = J, sigma_inv_cc
        self.F = g.zeros(0)
        self.Fnm1 = g.zeros(0)
        self.L = g.zeros(0)
        self.Tnm1 = self.T.copy(order="F")
        self.iters = 0

    def rk(self, name):
        return RK3[name][self.stage] if self.rk_stages == 3 else 1.0

    def sources(self, U):
        g = self.g
        rk_active = self.rk_stages == 3
        self.Fnm1 = self.F.copy(order="F")
        if rk_active:
            self.L = g.zeros(0)
        F = g.zeros(0)
        add, scale = self.advection
        if add:
            tF = [np.asfortranarray(O.interp(g, self.T, 0, 1 << d, d) * U[d]) for d in range(3)]
            t = O.div(g, tF[0], tF[1], tF[2], 0)
            t = t * (-1.0)
            F = F + t * scale
        add, scale = self.diffusion
        if add:
            t = lap_uniform_cc(g, self.T)
            if rk_active:
                self.L = np.asfortranarray(self.L + t * scale)
            else:
                F = F + t * scale
        U_CC = cell_center_velocity(g, U) if (self.KE_diffusion[0] or self.viscous_dissipation[0]) else None
        add, scale = self.KE_diffusion
        if add:
            F = F + ke_diffusion_term(g, U_CC) * scale
        add, scale = self.viscous_dissipation
        if add:
            F = F + viscous_dissipation_term(g, U_CC) * scale
        add, scale = self.joule_heating
        if add:
            F = F + joule_heating_term(g, self.J, self.sigma_inv_cc) * scale
        add, scale = self.volumetric_heating
        if add:
            F = F + volumetric_heating_profile(g, scale) * scale
        self.F = np.asfortranarray(F)

    def solve(self):
        if self.method == 1:
            self.T = np.asfortranarray(self.T + self.F * self.dt)
            return
        if self.method == 5:
            t = self.F * (self.dt * self.rk("gamma"))
            t = t + self.Fnm1 * (self.dt * self.rk("zeta"))
            t = t + self.L * (self.dt * self.rk("alpha"))
            rhs = np.asfortranarray(t + self.T)
            c = self.dt * self.rk("alpha") * self.coeff_implicit  # beta = alpha
        else:
            t = self.F * 1.5
            t = t + self.Fnm1 * (-0.5)
            if self.method == 4:
                t = t * (TWO_THIRDS * self.dt)
                t = t + self.T * FOUR_THIRDS
                rhs = np.asfortranarray(t + self.Tnm1 * NEG_ONE_THIRD)
                c = self.dt * TWO_THIRDS * self.coeff_implicit
            else:
                t = t * self.dt
                if self.method == 2:
                    self.T = np.asfortranarray(self.T + t)
                    return
                rhs = np.asfortranarray(t + self.T)
                c = self.dt * 1.0 * self.coeff_implicit
        self.Tnm1 = self.T.copy(order="F")
        self.iters, _, _, _ = O.pcg_nrg(self.g, self.Tbcs, self.all_neumann, self.k, c, self.T, rhs, *self.isp)

    def step(self, U):
        for self.stage in range(self.rk_stages if self.rk_stages == 3 else 1):
            self.sources(U)
            self.solve()
        self.stage = 0
