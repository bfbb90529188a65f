"""HIP-graph replay of the time step and the batched ensemble: a replayed
step equals the eagerly launched step bit for bit (every replay, with eager
work and synchronisations in between), and members of different sizes
advanced concurrently on several streams stay within 1e-10 of the oracle's
step, member by member."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import step_oracle as S

pytestmark = pytest.mark.gpu

ISP = (5, 1e-6, 1e-13, 1)


def state_arrays(mb):
    return [u.download() for u in mb.U] + [mb.p.download()] + [b.download() for b in mb.B]


def test_graph_replay_equals_eager_step(gpu):
    from moons_amd.ensemble import Member

    print("graph replay vs eager", flush=True)
    a = Member(12, 400.0, 10.0, 7, 5e-4, ISP)
    b = Member(12, 400.0, 10.0, 7, 5e-4, ISP)
    b.st.capture()
    for n in range(4):
        a.st.step(1)
        b.st.launch(1)
        gpu.synchronize()
        # eager work between replays
        junk = gpu.SF(a.m, "CC").fill(float(n))
        gpu.synchronize()
        for x, y in zip(state_arrays(a), state_arrays(b)):
            np.testing.assert_array_equal(x, y)
        del junk


def oracle_member(mb, dt):
    g = O.Grid(*mb.hs)
    LDC = [O.BCs.make(["Dirichlet"] * 6, [0, 0, 0, 1.0, 0, 0]), O.BCs.make(["Dirichlet"] * 6),
           O.BCs.make(["Dirichlet"] * 6)]
    PV = [O.BCs.make(["Neumann", "Neumann"] + ["Dirichlet"] * 4),
          O.BCs.make(["Dirichlet"] * 2 + ["Neumann", "Neumann"] + ["Dirichlet"] * 2),
          O.BCs.make(["Dirichlet"] * 4 + ["Neumann", "Neumann"])]
    U0 = [u.download() for u in mb.U]
    return S.MHDState(g, U0, LDC, O.BCs.make(["Neumann"] * 6), Re=mb.Re, dt=dt, alpha_U=1.0,
                      B=[g.zeros(1 << d) for d in range(3)], Bbcs=PV, phibcs=O.BCs.make(["Dirichlet"] * 6),
                      Rem=1.0, Ha=mb.Ha, B0=(0.0, 1.0, 0.0), isp_U=ISP, isp_p=ISP, isp_B=ISP, isp_phi=ISP)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_ensemble_members_match_oracle(gpu):
    from moons_amd.ensemble import Ensemble

    print("ensemble vs oracle", flush=True)
    dt = 5e-4
    ens = Ensemble(n=4, sizes=(8, 12), seed=3, dt=dt, isp=ISP, n_streams=2)
    orcs = [oracle_member(mb, dt) for mb in ens.members]
    for n in range(3):
        ens.step(1)
        for o in orcs:
            o.step()
        for q, (mb, o) in enumerate(zip(ens.members, orcs)):
            for d in range(3):
                assert rel(mb.U[d].download(), o.U[d]) <= 1e-10, (n, q, "U", d)
                assert rel(mb.B[d].download(), o.B[d]) <= 1e-10, (n, q, "B", d)
            assert rel(mb.p.download()[1:-1, 1:-1, 1:-1], o.p[1:-1, 1:-1, 1:-1]) <= 1e-10, (n, q, "p")
