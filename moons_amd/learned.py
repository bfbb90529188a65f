"""Training data and set-up for the learned PPE preconditioner.

A sample is (input, target) on one cell-centred grid of n0 x n1 x n2 cells:
  input  4 channels: g = f vol^(2/3) / s, and log(dh_d / vol^(1/3)), d = x, y, z
  target e / RMS(e), where lap(e) = f (the PPE's Laplacian with its BCs; the
         constant removed for all-Neumann problems); s = RMS of f vol^(2/3).
         The flexible PCG is invariant to a constant factor on z, so only the
         shape of e is learned; unit-RMS targets keep the training well scaled
as FP32 arrays with the padded column-major shape of a cell-centred field.
The device preconditioner forms the same input from r/vol and returns
z = s N(input).  Right-hand sides come from solver runs (the pressure
right-hand sides of MHD_step) or from seeded random fields; targets are
solved on the device with the FP32-multigrid flexible PCG to tol 1e-10.
"""
from __future__ import annotations

import numpy as np

from . import core as M


def geometry(hs):
    """vol^(2/3) and the three log aspect channels on the padded CC grid."""
    dh = [np.diff(np.asarray(h, dtype=np.float64)) for h in hs]  # padded cell widths
    DX, DY, DZ = np.meshgrid(dh[0], dh[1], dh[2], indexing="ij")
    vol = DX * DY * DZ
    h = np.cbrt(vol)
    return h * h, [np.log(DX / h), np.log(DY / h), np.log(DZ / h)]


def make_sample(f, e, hs):
    """(input (4, s0, s1, s2), target (s0, s1, s2)) from f and e (padded CC arrays)."""
    v23, geom = geometry(hs)
    g = f * v23
    inner = (slice(1, -1),) * 3
    s = float(np.sqrt(np.mean(g[inner] ** 2)))
    s = s if s > 0 else 1.0
    x = np.zeros((4,) + f.shape, dtype=np.float32)
    x[0][inner] = (g / s)[inner]
    for d in range(3):
        x[1 + d][inner] = geom[d][inner]
    t = np.zeros(f.shape, dtype=np.float32)
    et = e[inner]
    rms = float(np.sqrt(np.mean(et * et)))
    t[inner] = et / (rms if rms > 0 else 1.0)
    return x, t


def solve_targets(m, types, rhs_list, tol=1e-10):
    """e with lap(e) = f for each f (padded CC arrays), on the device (FP32 MG in FP64 FPCG)."""
    alln = all(t in ("Neumann", "Periodic") for t in types)
    X = M.SF(m, "CC").set_BCs(types, None, all_Neumann=alln)
    P = M.PCG_solver_SF(m, X, M.iter_solver_params(500, tol, 1e-300, 1), prec="MG_FP32")
    B = M.SF(m, "CC")
    out = []
    for f in rhs_list:
        X.fill(0.0)
        B.upload(f)
        P.solve(X, B)
        e = X.download()
        if alln:
            e = e - e[1:-1, 1:-1, 1:-1].mean()
        out.append(e)
    return out


def random_rhs(shape, rng, n_modes=6):
    """Seeded random right-hand side: a few random Fourier modes plus white noise, mean-free."""
    n0, n1, n2 = (s - 2 for s in shape)
    x, y, z = np.meshgrid(*[(np.arange(n) + 0.5) / n for n in (n0, n1, n2)], indexing="ij")
    f = np.zeros((n0, n1, n2))
    for _ in range(n_modes):
        k = rng.integers(1, 6, size=3)
        ph = rng.random(3) * 2 * np.pi
        f += rng.standard_normal() * np.cos(np.pi * k[0] * x + ph[0]) * np.cos(np.pi * k[1] * y + ph[1]) * \
            np.cos(np.pi * k[2] * z + ph[2])
    f += 0.3 * rng.standard_normal(f.shape) * rng.random()
    f -= f.mean()
    out = np.zeros(shape)
    out[1:-1, 1:-1, 1:-1] = f
    return np.asfortranarray(out)


def random_pairs(hs, types, n, seed=0):
    """n samples with random right-hand sides on the mesh of node arrays hs."""
    m = M.mesh(*hs)
    shape = tuple(len(h) - 1 for h in hs)
    rng = np.random.default_rng(seed)
    rhs = [random_rhs(shape, rng) for _ in range(n)]
    es = solve_targets(m, types, rhs)
    return [make_sample(f, e, hs) for f, e in zip(rhs, es)]


def run_pairs(N, n_runs, steps, seed=0, every=1, Ha_max=0.0):
    """Samples from lid-driven-cavity runs on an N^3 uniform unit cube: each run
    has a seeded Re in [100, 1000] (and Ha in [0, Ha_max] with the low-Rem
    induction when Ha_max > 0), 1e-3 seeded noise in U; the pressure right-hand
    side of every `every`-th step is solved accurately for the target."""
    rng = np.random.default_rng(seed)
    hs = [M.nodes_uniform(0.0, 1.0, N)] * 3
    m = M.mesh(*hs)
    ubc = [(["Dirichlet"] * 6, [0, 0, 0, 1.0, 0, 0]), (["Dirichlet"] * 6, None), (["Dirichlet"] * 6, None)]
    pv = [(["Neumann", "Neumann"] + ["Dirichlet"] * 4, None),
          (["Dirichlet"] * 2 + ["Neumann", "Neumann"] + ["Dirichlet"] * 2, None),
          (["Dirichlet"] * 4 + ["Neumann", "Neumann"], None)]
    rhs = []
    for _ in range(n_runs):
        Re = float(rng.uniform(100, 1000))
        Ha = float(rng.uniform(0, Ha_max)) if Ha_max > 0 else 0.0
        U = []
        for d in range(3):
            f = M.SF(m, 1 << d)
            a = (rng.random(f.shape) * 2 - 1) * 1e-3
            f.upload(np.asfortranarray(a)).set_BCs(*ubc[d])
            M.apply_BCs(f)
            U.append(f)
        p = M.SF(m, "CC").fill(0.0).set_BCs(["Neumann"] * 6, None, all_Neumann=True)
        kw = dict(Re=Re, dt=1e-3, alpha_U=1.0, isp_p=(200, 1e-8, 1e-300, 1), isp_U=(50, 1e-8, 1e-300, 1))
        if Ha_max > 0:
            B = [M.SF(m, 1 << d).fill(0.0).set_BCs(*pv[d]) for d in range(3)]
            phi = M.SF(m, "CC").fill(0.0).set_BCs(["Dirichlet"] * 6)
            kw.update(B=B, phi=phi, Ha=Ha, Rem=1.0, isp_B=(50, 1e-8, 1e-300, 1), isp_phi=(200, 1e-8, 1e-300, 1))
        st = M.MHD_step(m, U, p, **kw)
        for s in range(steps):
            st.step(1)
            if s % every == 0:
                rhs.append(np.asfortranarray(st.field("ppe_rhs")[0].download()))
    types = ["Neumann"] * 6
    es = solve_targets(m, types, rhs)
    return [make_sample(f, e, hs) for f, e in zip(rhs, es)]
