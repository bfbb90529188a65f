"""GPU parity of the time step at the benchmark configurations' own sizes,
against the oracle's step, step by step:

* uniform 32^3 hydrodynamic lid-driven cavity, Re 400, dt 1e-3, methods 4
  and 6, PPE by Jacobi-PCG to 1e-8 (iter_max 10^4), 50 steps from rest plus
  seeded noise of amplitude 1e-3;
* uniform 64^3 MHD cavity, Re 400, Ha 10, Rem 1 (low Rem, Btot = B0),
  B0 = (0, 1, 0), pseudo-vacuum B walls, phi Dirichlet, dt 5e-4, 10 steps;
  PPE to 1e-8 and phi cleaning to 1e-12.

After every step U, p and F (and B, J, J x B) agree within 1e-10 relative L2
and every solve takes the same number of iterations.  div B is enforced by
the phi cleaning (clean_divergence.f90:37-59): the implicit induction solve
(method 6) leaves div B at its PCG tolerance, and div B after cleaning is the
phi solve's residual, so phi is solved to 1e-12; every step of the MHD run
asserts max|div B| h / max|B| <= 1e-12 on the device's field."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import step_oracle as S

pytestmark = pytest.mark.gpu

LDC = [(["Dirichlet"] * 6, [0, 0, 0, 1.0, 0, 0]), (["Dirichlet"] * 6, None), (["Dirichlet"] * 6, None)]
PV = [(["Neumann", "Neumann"] + ["Dirichlet"] * 4, None),
      (["Dirichlet"] * 2 + ["Neumann", "Neumann"] + ["Dirichlet"] * 2, None),
      (["Dirichlet"] * 4 + ["Neumann", "Neumann"], None)]
PRESET = (5, 1e-6, 1e-13, 1)  # the reference's ISP for U and B (sim_params_init_solver_config.f90)
PPE = (10000, 1e-8, 1e-300, 1)
PHI = (10000, 1e-12, 1e-300, 1)


def rel_l2(a, b):
    den = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b)) / (den if den > 0 else 1.0)


def seeded_noise(g, seed, amp=1e-3):
    """SplitMix64 noise on the interior faces of each component (seed 10 s + d), then the lid BCs."""
    out = []
    for d in range(3):
        shape = g.shape(1 << d)
        u = np.zeros(shape, order="F")
        core = tuple(slice(2, s - 2) if (1 << d) >> e & 1 else slice(1, s - 1) for e, s in enumerate(shape))
        n = int(np.prod([c.stop - c.start for c in core]))
        u[core] = O.splitmix(n, 10 * seed + d, amp).reshape([c.stop - c.start for c in core], order="F")
        u = np.asfortranarray(u)
        O.apply_bcs(g, u, 1 << d, O.BCs.make(*LDC[d]))
        out.append(u)
    return out


def max_div_B(g, B, N):
    dv = O.div(g, B[0], B[1], B[2], 0)[1:-1, 1:-1, 1:-1]
    bmax = max(float(np.abs(b).max()) for b in B)
    return float(np.abs(dv).max()) * (1.0 / N) / bmax if bmax > 0 else 0.0


def run_config(gpu, N, nsteps, mhd, alpha, dt, seed):
    hs = [O.uniform_nodes(0.0, 1.0, N)] * 3
    g, m = O.Grid(*hs), gpu.mesh(*hs)
    U0 = seeded_noise(g, seed)
    Ug = [gpu.SF(m, 1 << d).upload(U0[d]).set_BCs(*LDC[d]) for d in range(3)]
    pg = gpu.SF(m, "CC").fill(0.0).set_BCs(["Neumann"] * 6, None, all_Neumann=True)
    kw = dict(Re=400.0, dt=dt, alpha_U=alpha, isp_U=PRESET, isp_p=PPE)
    okw = dict(kw)
    if mhd:
        Bg = [gpu.SF(m, 1 << d).fill(0.0).set_BCs(*PV[d]) for d in range(3)]
        phig = gpu.SF(m, "CC").fill(0.0).set_BCs(["Dirichlet"] * 6)
        common = dict(Rem=1.0, Ha=10.0, B0=(0.0, 1.0, 0.0), isp_B=PRESET, isp_phi=PHI)
        kw.update(B=Bg, phi=phig, **common)
        okw.update(B=[g.zeros(1 << d) for d in range(3)], Bbcs=[O.BCs.make(*b) for b in PV],
                   phibcs=O.BCs.make(["Dirichlet"] * 6), **common)
    st = gpu.MHD_step(m, Ug, pg, **kw)
    ost = S.MHDState(g, U0, [O.BCs.make(*u) for u in LDC], O.BCs.make(["Neumann"] * 6), **okw)
    worst = 0.0
    for n in range(nsteps):
        info = st.step(1)
        ost.step()
        its = info["iterations"]
        assert its["p"] == ost.iters["p"], (n, its, ost.iters)
        if alpha:
            assert its["U"] == ost.iters["U"], (n, its, ost.iters)
        errs = {}
        for d in range(3):
            errs[f"U{d}"] = rel_l2(Ug[d].download(), ost.U[d])
            errs[f"F{d}"] = rel_l2(st.field("F")[d].download(), ost.F[d])
        errs["p"] = rel_l2(pg.download()[1:-1, 1:-1, 1:-1], ost.p[1:-1, 1:-1, 1:-1])
        if mhd:
            assert its["B"] == ost.iters["B"] and its["phi"] == ost.iters["phi"], (n, its, ost.iters)
            Bd = [kw["B"][d].download() for d in range(3)]
            for d in range(3):
                errs[f"B{d}"] = rel_l2(Bd[d], ost.B[d])
                errs[f"J{d}"] = rel_l2(st.field("J")[d].download(), ost.J[d])
                errs[f"JxB{d}"] = rel_l2(st.field("jCrossB")[d].download(), ost.jCrossB[d])
            divb = max_div_B(g, Bd, N)
            assert divb <= 1e-12, (n, divb)
        bad = {k: v for k, v in errs.items() if not v <= 1e-10}
        assert not bad, (n, bad)
        worst = max(worst, max(errs.values()))
        if n % 10 == 0 or n == nsteps - 1:
            extra = f" max|divB|h/max|B| {divb:.2e}" if mhd else ""
            print(f"  step {n}: iterations {its}, worst rel L2 so far {worst:.2e}{extra}", flush=True)
    return worst


@pytest.mark.timeout(900)
@pytest.mark.parametrize("method", [4, 6])
def test_c1_hydro_ldc_32(gpu, method):
    print(f"C1 32^3 LDC Re 400 method {method}, 50 steps, PPE to 1e-8", flush=True)
    run_config(gpu, 32, 50, False, 0.0 if method == 4 else 1.0, 1e-3, 1)


@pytest.mark.timeout(900)
def test_c2_mhd_ldc_64(gpu):
    print("C2 64^3 MHD LDC Re 400 Ha 10 method 6, 10 steps, PPE to 1e-8, phi to 1e-12", flush=True)
    run_config(gpu, 64, 10, True, 1.0, 5e-4, 2)
