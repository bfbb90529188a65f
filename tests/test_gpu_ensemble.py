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


@pytest.mark.timeout(600)
def test_ensemble_uneven_sizes_match_oracle(gpu):
    """The batched ensemble's sizes: one member of each of 32^3, 48^3, 64^3
    and 96^3 (Re in [100, 1000], Ha in [0, 50], seeded noise), advanced by
    graph replays on two streams for 3 steps; every member's U, B and p stay
    within 1e-10 of the oracle's step (reported per member)."""
    from moons_amd.ensemble import Ensemble

    print("ensemble of 32, 48, 64, 96 vs oracle", flush=True)
    dt = 5e-4
    ens = Ensemble(seed=5, dt=dt, isp=ISP, n_streams=2, member_sizes=(32, 48, 64, 96))
    orcs = [oracle_member(mb, dt) for mb in ens.members]
    for n in range(3):
        ens.step(1)
        worst = []
        for q, (mb, o) in enumerate(zip(ens.members, orcs)):
            o.step()
            e = max([rel(mb.U[d].download(), o.U[d]) for d in range(3)] +
                    [rel(mb.B[d].download(), o.B[d]) for d in range(3)] +
                    [rel(mb.p.download()[1:-1, 1:-1, 1:-1], o.p[1:-1, 1:-1, 1:-1])])
            worst.append(e)
            assert e <= 1e-10, (n, mb.N, e)
        print(f"  step {n}: worst rel L2 per member " +
              ", ".join(f"{mb.N}^3 (Re {mb.Re:.0f}, Ha {mb.Ha:.1f}) {e:.1e}" for mb, e in zip(ens.members, worst)),
              flush=True)


def test_64_graphs_on_four_streams_equal_eager(gpu):
    """The ensemble's launch pattern at its member count: 64 captured steps
    (8^3, 12^3 and 16^3 members) replayed over 4 streams for 3 steps equal the
    same members stepped eagerly, bit for bit."""
    from moons_amd.ensemble import Ensemble

    print("64 graphs on 4 streams vs eager", flush=True)
    kw = dict(n=64, sizes=(8, 12, 16), seed=11, dt=5e-4, isp=ISP, n_streams=4)
    ens = Ensemble(**kw)
    ref = Ensemble(capture=False, **kw)
    assert [mb.N for mb in ens.members] == [mb.N for mb in ref.members]
    for n in range(3):
        ens.step(1)
        for mb in ref.members:
            mb.st.step(1)
        gpu.synchronize()
        for q, (a, b) in enumerate(zip(ens.members, ref.members)):
            for x, y in zip(state_arrays(a), state_arrays(b)):
                assert np.array_equal(x, y), (n, q, a.N)
