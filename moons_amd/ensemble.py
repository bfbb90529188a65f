"""Batched ensemble of independent MHD lid-driven-cavity instances.

Each member is an MHD_step state on its own uniform N^3 mesh (N drawn from
`sizes`), with seeded Re and Ha, the lid-driven cavity velocity BCs, the
insulating (pseudo-vacuum) B BCs, B0 = (0, 1, 0) and 1e-3 seeded velocity
noise.  Every member's time step is captured once into a HIP graph; a step of
the ensemble replays each member's graph on one of a few streams, so the
launch-bound small members run concurrently.
"""
from __future__ import annotations

import numpy as np

from . import core as M

LDC_U = [(["Dirichlet"] * 6, [0, 0, 0, 1.0, 0, 0]), (["Dirichlet"] * 6, None), (["Dirichlet"] * 6, None)]
PSEUDO_VACUUM = [(["Neumann", "Neumann"] + ["Dirichlet"] * 4, None),
                 (["Dirichlet"] * 2 + ["Neumann", "Neumann"] + ["Dirichlet"] * 2, None),
                 (["Dirichlet"] * 4 + ["Neumann", "Neumann"], None)]


def noise_velocity(m, seed, amp=1e-3):
    rng = np.random.default_rng(seed)
    U = []
    for d in range(3):
        f = M.SF(m, 1 << d)
        f.upload(np.asfortranarray(amp * (2.0 * rng.random(f.shape) - 1.0))).set_BCs(*LDC_U[d])
        M.apply_BCs(f)
        U.append(f)
    return U


class Member:
    def __init__(self, N, Re, Ha, seed, dt, isp, induction=True):
        self.N, self.Re, self.Ha, self.seed = N, Re, Ha, seed
        hs = [M.nodes_uniform(0.0, 1.0, N)] * 3
        self.hs = hs
        self.m = M.mesh(*hs)
        self.U = noise_velocity(self.m, seed)
        self.p = M.SF(self.m, "CC").fill(0.0).set_BCs(["Neumann"] * 6, None, all_Neumann=True)
        kw = dict(Re=Re, dt=dt, alpha_U=1.0, isp_U=isp, isp_p=isp)
        self.B = self.phi = None
        if induction:
            self.B = [M.SF(self.m, 1 << d).fill(0.0).set_BCs(*PSEUDO_VACUUM[d]) for d in range(3)]
            self.phi = M.SF(self.m, "CC").fill(0.0).set_BCs(["Dirichlet"] * 6)
            kw.update(B=self.B, phi=self.phi, Ha=Ha, Rem=1.0, B0=(0.0, 1.0, 0.0), isp_B=isp, isp_phi=isp)
        self.st = M.MHD_step(self.m, self.U, self.p, **kw)


class Ensemble:
    def __init__(self, n=64, sizes=(32, 48, 64, 96), Re=(100.0, 1000.0), Ha=(0.0, 50.0), seed=5, dt=5e-4,
                 isp=(5, 1e-6, 1e-13, 1), n_streams=4, capture=True, member_sizes=None):
        """n members with N drawn uniformly from sizes (or exactly member_sizes),
        Re and Ha uniform in their ranges, all seeded by seed."""
        rng = np.random.default_rng(seed)
        self.members = []
        if member_sizes is not None:
            n = len(member_sizes)
        for q in range(n):
            N = int(member_sizes[q]) if member_sizes is not None else int(rng.choice(sizes))
            self.members.append(Member(N, float(rng.uniform(*Re)), float(rng.uniform(*Ha)), seed * 1000 + q, dt, isp))
        self.streams = [M.Stream() for _ in range(n_streams)]
        if capture:
            for mb in self.members:
                mb.st.capture()

    @property
    def cells(self):
        return sum(mb.N ** 3 for mb in self.members)

    def step(self, n=1):
        """n time steps of every member: graph replays over the streams."""
        K = len(self.streams)
        for _ in range(n):
            for q, mb in enumerate(self.members):
                mb.st.launch(1, self.streams[q % K])
        for s in self.streams:
            s.synchronize()
