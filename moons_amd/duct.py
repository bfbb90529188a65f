"""The rectangular-duct MHD flow (Hartmann/Shercliff type) on the device.

The reference's MHD_Shercliff preset
(code/handwritten/sim_params/sim_params_default_MHD_Shercliff/*.f90) with its
periodic flow direction, Roberts-clustered walls and thin conducting walls:

* mesh: x uniform and periodic, y and z Roberts (robertsBoth, beta) on
  [-1, 1] (define_mesh_MHD_Shercliff_plane, code/handwritten/sim_params/mesh_bank.f90:176-190,
  with n cells along x);
* U: no-slip walls, x periodic (init_U_BCs preset 0 + make_periodic); p: Neumann
  walls, x periodic, all_Neumann (init_P_BCs preset 0);
* B: thin conducting walls (periodic_duct_thin_wall,
  code/handwritten/solvers/induction/init_B_BCs.f90:124-175): the wall-normal
  component Neumann (pseudo-vacuum), the tangential ones Robin with the wall
  conductance ratio c_w, x periodic; phi: Dirichlet walls, x periodic;
* B0 = (0, 0, 1) (uniform_B0_dir = 3), finite Rem (Btot = B0 + B), the flow
  driven along x by the mean pressure gradient (mpg_dir = 1, scale 1,
  sim_params_init_logicals.f90:62-65, sim_params_init_flow_config.f90:54);
* backward-Euler diffusion of U and B (method 6); the reference's 5-iteration
  solver settings for U and B, the PPE to 1e-8 and phi cleaning to 1e-12.

Velocity noise (seeded SplitMix64, amplitude 1e-3) on the interior faces makes
the first steps non-trivial.  `prec_p` picks the PPE solver and `prec_phi`
the phi cleaning's: "Lap_SF" (the FP64 Jacobi-PCG drop-in) or "MG_FP32" (the
mixed-precision multigrid).
"""
from __future__ import annotations

import numpy as np

from . import core as M

PRESET_ISP = (5, 1e-6, 1e-13, 1)  # sim_params_init_solver_config.f90:46-52
PPE_ISP = (10000, 1e-8, 1e-300, 1)
PHI_ISP = (10000, 1e-12, 1e-300, 1)

PER_X = ["Periodic", "Periodic"]


def bc_sets(c_w: float):
    """(types, values, c_w) of U (three components), p, B (three) and phi."""
    cw = [c_w] * 6
    U = [(PER_X + ["Dirichlet"] * 4, None, None)] * 3
    p = (PER_X + ["Neumann"] * 4, None, None)
    B = [(PER_X + ["Robin"] * 4, None, cw),
         (PER_X + ["Neumann", "Neumann", "Robin", "Robin"], None, cw),
         (PER_X + ["Robin", "Robin", "Neumann", "Neumann"], None, cw)]
    phi = (PER_X + ["Dirichlet"] * 4, None, None)
    return U, p, B, phi


def splitmix(n: int, seed: int, amp: float) -> np.ndarray:
    """SplitMix64 counter form -> uniform [-amp, amp)."""
    q = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + q * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return amp * (2.0 * u - 1.0)


def duct_nodes(n, beta=1.05):
    """(hx, hy, hz) with ghosts: n = (nx, ny, nz) cells."""
    return (M.nodes_uniform(-1.0, 1.0, n[0]), M.nodes_roberts_both(-1.0, 1.0, n[1], beta),
            M.nodes_roberts_both(-1.0, 1.0, n[2], beta))


def noise_faces(shapes, seed, amp=1e-3):
    """Seeded noise on the interior faces of each component (zero elsewhere)."""
    out = []
    for d, shape in enumerate(shapes):
        u = np.zeros(shape, order="F")
        core = tuple(slice(2, s - 2) if (1 << d) >> e & 1 else slice(1, s - 1) for e, s in enumerate(shape))
        dims = [c.stop - c.start for c in core]
        u[core] = splitmix(int(np.prod(dims)), 10 * seed + d, amp).reshape(dims, order="F")
        out.append(np.asfortranarray(u))
    return out


class Duct:
    """Device state of the duct flow: mesh, fields and the MHD step."""

    def __init__(self, n=(128, 128, 128), beta=1.05, Re=1000.0, Ha=100.0, Rem=1.0, c_w=0.01, dt=1e-3,
                 prec_p="MG_FP32", seed=3, isp_p=PPE_ISP, isp_phi=PHI_ISP, isp_U=PRESET_ISP, isp_B=PRESET_ISP,
                 prec_phi="Lap_SF"):
        self.n, self.beta, self.Re, self.Ha, self.Rem, self.c_w, self.dt = tuple(n), beta, Re, Ha, Rem, c_w, dt
        self.isp = {"U": isp_U, "p": isp_p, "B": isp_B, "phi": isp_phi}
        self.hs = duct_nodes(self.n, beta)
        self.m = M.mesh(*self.hs)
        Ub, pb, Bb, phib = bc_sets(c_w)
        self.U0 = noise_faces([self.m.shape(1 << d) for d in range(3)], seed)
        self.U = []
        for d in range(3):
            f = M.SF(self.m, 1 << d).upload(self.U0[d]).set_BCs(Ub[d][0], Ub[d][1])
            M.apply_BCs(f)
            self.U.append(f)
        self.p = M.SF(self.m, "CC").fill(0.0).set_BCs(pb[0], None, all_Neumann=True)
        self.B = [M.SF(self.m, 1 << d).fill(0.0).set_BCs(Bb[d][0], None, c_w=Bb[d][2]) for d in range(3)]
        self.phi = M.SF(self.m, "CC").fill(0.0).set_BCs(phib[0])
        self.st = M.MHD_step(self.m, self.U, self.p, B=self.B, phi=self.phi, Re=Re, Ha=Ha, Rem=Rem, dt=dt,
                             alpha_U=1.0, B0=(0.0, 0.0, 1.0), isp_U=isp_U, isp_p=isp_p, isp_B=isp_B,
                             isp_phi=isp_phi, prec_p=prec_p, mpg_dir=1, mpg_scale=1.0, finite_Rem=True,
                             prec_phi=prec_phi)

    def step(self, k: int = 1):
        return self.st.step(k)

    def fields(self):
        """Host copies: U, p, B (three faces each but p) and J (three edges)."""
        return {"U": [u.download() for u in self.U], "p": self.p.download(),
                "B": [b.download() for b in self.B], "J": [j.download() for j in self.st.field("J")[:3]]}
