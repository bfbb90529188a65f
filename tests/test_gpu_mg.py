"""GPU FP32 multigrid: one V-cycle against the FP64 NumPy restatement on the
same residual (FP32 tolerance), and the mixed-precision flexible PCG against
the FP64 oracle's Jacobi-PCG solution (after removing the free constant of the
singular Neumann problem) at the 1e-6 bound stated for mixed-precision fields."""
import ctypes as C

import numpy as np
import pytest

from oracle import mg_oracle as MGO
from oracle import oracle as O

pytestmark = pytest.mark.gpu

NAMES = {1: "Dirichlet", 2: "Neumann", 3: "Periodic"}
CASES = {
    "cube16_neumann": (16, None, [2] * 6),
    "duct24_roberts": (24, 1.05, [3, 3, 2, 2, 2, 2]),
    "dirichlet_mix20": (20, 1.2, [1, 2, 1, 1, 2, 1]),
    # zebra y- and z-lines on the finest level (24^3 smooths by points)
    "duct48_lines": (48, 1.05, [3, 3, 2, 2, 2, 2]),
    # z periodic: y-lines only
    "ylines48_zperiodic": (48, 1.1, [2, 1, 1, 2, 3, 3]),
}


def meshes(gpu, N, beta):
    hx = O.uniform_nodes(0.0, 1.0, N)
    hy = hx if beta is None else O.roberts_both_nodes(0.0, 1.0, N, beta)
    return [hx, hy, hy], O.Grid(hx, hy, hy), gpu.mesh(hx, hy, hy)


def mg_solver(gpu, m, x, iters=200, tol=1e-8):
    P = gpu.PCG_solver_SF(m, x, gpu.iter_solver_params(iters, tol, 1e-300, 1), prec="MG_FP32")
    return P


@pytest.mark.parametrize("case", list(CASES))
def test_vcycle_matches_restatement(gpu, case):
    N, beta, bct = CASES[case]
    hs, g, m = meshes(gpu, N, beta)
    x = gpu.SF(m, "CC").set_BCs([NAMES[t] for t in bct], None, all_Neumann=(1 not in bct))
    P = mg_solver(gpu, m, x)
    rng = np.random.default_rng(3)
    r = np.zeros(g.shape(0), order="F")
    r[1:-1, 1:-1, 1:-1] = rng.standard_normal(tuple(n - 2 for n in g.shape(0)))
    R = gpu.SF(m, "CC").upload(r)
    Z = gpu.SF(m, "CC")
    P.apply_prec(R, Z)
    z = Z.download()
    levels = MGO.build(hs, bct)
    print(case, "line directions per level", [L.line_dirs for L in levels], flush=True)
    vol = O.volume(g, 0)
    f = np.zeros(g.shape(0))
    f[1:-1, 1:-1, 1:-1] = r[1:-1, 1:-1, 1:-1] / vol[1:-1, 1:-1, 1:-1]
    zr = MGO.vcycle(levels, f, bct)
    zi, zri = z[1:-1, 1:-1, 1:-1], zr[1:-1, 1:-1, 1:-1]
    err = np.linalg.norm(zi - zri) / np.linalg.norm(zri)
    assert err < 2e-5, err
    assert P.mg_info()[0] == len(levels)


@pytest.mark.parametrize("nu", [(1, 0, 16), (2, 1, 16), (3, 2, 8), (1, 3, 4)])
@pytest.mark.parametrize("case", ["duct48_lines", "ylines48_zperiodic", "dirichlet_mix20"])
def test_vcycle_sweep_counts_match_restatement(gpu, case, nu):
    """Sweep counts other than 2 + 2 (moons_pcg_set_mg) against the
    restatement, whose prolongation corrects every fine cell.  With no
    post-smoothing the device does too; otherwise it corrects only the cells
    that the first post-smoothing half-sweep reads (z-line, y-line or
    red-black colour 0, on the grid levels and in the LDS tail: the 48^3 duct,
    the y-lines with z periodic, and the 20^3 hierarchy held in the tail)."""
    nu1, nu2, cs = nu
    N, beta, bct = CASES[case]
    print(f"sweep counts {case} {nu}", flush=True)
    hs, g, m = meshes(gpu, N, beta)
    x = gpu.SF(m, "CC").set_BCs([NAMES[t] for t in bct], None, all_Neumann=(1 not in bct))
    P = mg_solver(gpu, m, x)
    P.set_MG(16, nu1, nu2, cs)
    rng = np.random.default_rng(31)
    r = np.zeros(g.shape(0), order="F")
    r[1:-1, 1:-1, 1:-1] = rng.standard_normal(tuple(n - 2 for n in g.shape(0)))
    Z = gpu.SF(m, "CC")
    P.apply_prec(gpu.SF(m, "CC").upload(r), Z)
    z = Z.download()
    levels = MGO.build(hs, bct)
    vol = O.volume(g, 0)
    f = np.zeros(g.shape(0))
    f[1:-1, 1:-1, 1:-1] = r[1:-1, 1:-1, 1:-1] / vol[1:-1, 1:-1, 1:-1]
    zr = MGO.vcycle(levels, f, bct, nu1=nu1, nu2=nu2, coarse_sweeps=cs)
    zi, zri = z[1:-1, 1:-1, 1:-1], zr[1:-1, 1:-1, 1:-1]
    err = np.linalg.norm(zi - zri) / np.linalg.norm(zri)
    assert err < 2e-5, err


@pytest.mark.parametrize("level,cycles", [(0, 2), (1, 3)])
def test_gamma_cycle_matches_restatement(gpu, level, cycles):
    """The V-cycle with the part from `level` down cycled `cycles` times per
    visit, on the 48^3 stretched duct (zebra lines on the finest level),
    against the restatement; cycles = 1 equals the plain V-cycle bit for bit."""
    N, beta, bct = CASES["duct48_lines"]
    print(f"gamma-cycle level {level} x{cycles}", flush=True)
    hs, g, m = meshes(gpu, N, beta)
    x = gpu.SF(m, "CC").set_BCs([NAMES[t] for t in bct], None, all_Neumann=(1 not in bct))
    P = mg_solver(gpu, m, x)
    rng = np.random.default_rng(4)
    r = np.zeros(g.shape(0), order="F")
    r[1:-1, 1:-1, 1:-1] = rng.standard_normal(tuple(n - 2 for n in g.shape(0)))
    R = gpu.SF(m, "CC").upload(r)
    Z0, Z1, Z = gpu.SF(m, "CC"), gpu.SF(m, "CC"), gpu.SF(m, "CC")
    P.apply_prec(R, Z0)
    P.set_MG_gamma(level, 1)
    P.apply_prec(R, Z1)
    np.testing.assert_array_equal(Z1.download(), Z0.download())
    P.set_MG_gamma(level, cycles)
    P.apply_prec(R, Z)
    z = Z.download()
    levels = MGO.build(hs, bct)
    vol = O.volume(g, 0)
    f = np.zeros(g.shape(0))
    f[1:-1, 1:-1, 1:-1] = r[1:-1, 1:-1, 1:-1] / vol[1:-1, 1:-1, 1:-1]
    zr = MGO.vcycle(levels, f, bct, gamma_level=level, gamma_cycles=cycles)
    zi, zri = z[1:-1, 1:-1, 1:-1], zr[1:-1, 1:-1, 1:-1]
    err = np.linalg.norm(zi - zri) / np.linalg.norm(zri)
    assert err < 2e-5, err
    # the repeated coarse part changes the result
    assert np.linalg.norm(zi - Z0.download()[1:-1, 1:-1, 1:-1]) > 1e-6 * np.linalg.norm(zri)
    # a level of the single-workgroup coarse end is refused
    with pytest.raises(gpu.MoonsError, match="grid-launched"):
        P.set_MG_gamma(P.mg_info()[0] - 1, 2)


@pytest.mark.parametrize("bct", [[3, 3, 2, 2, 2, 2], [2, 1, 3, 3, 2, 2]])
def test_vcycle_half_sweep_levels_match_restatement(gpu, bct):
    """A mesh whose finest level has 80 x 64 x 64 cells (above the fused-sweep
    size): one V-cycle against the restatement.  Periodic along x with y and z
    stretched, the two finest levels are smoothed by zebra y- and z-lines;
    periodic along z with y uniform, the finest level by the in-place
    red-black half-sweeps and the next by the fused whole sweeps."""
    print(f"vcycle 80x64x64 {bct}", flush=True)
    hx = O.uniform_nodes(0.0, 1.25, 80)
    hy = O.roberts_both_nodes(0.0, 1.0, 64, 1.05) if bct[2] != 3 else O.uniform_nodes(0.0, 1.0, 64)
    hz = O.roberts_both_nodes(0.0, 1.0, 64, 1.1)
    hs = [hx, hy, hz]
    g, m = O.Grid(*hs), gpu.mesh(*hs)
    x = gpu.SF(m, "CC").set_BCs([NAMES[t] for t in bct], None, all_Neumann=(1 not in bct))
    P = mg_solver(gpu, m, x)
    rng = np.random.default_rng(8)
    r = np.zeros(g.shape(0), order="F")
    r[1:-1, 1:-1, 1:-1] = rng.standard_normal(tuple(n - 2 for n in g.shape(0)))
    R = gpu.SF(m, "CC").upload(r)
    Z = gpu.SF(m, "CC")
    P.apply_prec(R, Z)
    z = Z.download()
    levels = MGO.build(hs, bct)
    assert np.prod(levels[0].n) > 64 ** 3 >= np.prod(levels[1].n)
    vol = O.volume(g, 0)
    f = np.zeros(g.shape(0))
    f[1:-1, 1:-1, 1:-1] = r[1:-1, 1:-1, 1:-1] / vol[1:-1, 1:-1, 1:-1]
    zr = MGO.vcycle(levels, f, bct)
    zi, zri = z[1:-1, 1:-1, 1:-1], zr[1:-1, 1:-1, 1:-1]
    err = np.linalg.norm(zi - zri) / np.linalg.norm(zri)
    assert err < 2e-5, err


RAGGED = {
    # y-lines of 100 cells in 128-cell slots, z-lines of 36 in 64; z Dirichlet at one wall
    "x40_y100_z36": ((40, 100, 36), (1.05, 1.1), [3, 3, 2, 2, 1, 2]),
    # y- and z-lines of 24 cells (32-cell slots)
    "x64_y24_z24": ((64, 24, 24), (1.02, 1.05), [2, 2, 2, 1, 2, 2]),
    # y-lines of 40 (64-cell slots), z-lines of 12 (16-cell slots)
    "x128_y40_z12": ((128, 40, 12), (1.05, 1.02), [3, 3, 1, 2, 2, 2]),
}


@pytest.mark.parametrize("case", list(RAGGED))
def test_vcycle_ragged_lines_match_restatement(gpu, case):
    """Zebra lines whose length is not the slot size: every line-kernel
    instantiation (16, 32, 64 and 128-cell slots), padded rows, and y- and
    z-lines of different lengths on one level, against the restatement."""
    (nx, ny, nz), (by, bz), bct = RAGGED[case]
    print(f"vcycle ragged lines {case}", flush=True)
    hs = [O.uniform_nodes(0.0, 1.0, nx), O.roberts_both_nodes(0.0, 1.0, ny, by),
          O.roberts_both_nodes(0.0, 1.0, nz, bz)]
    g, m = O.Grid(*hs), gpu.mesh(*hs)
    levels = MGO.build(hs, bct)
    assert levels[0].line_dirs == [1, 2]
    x = gpu.SF(m, "CC").set_BCs([NAMES[t] for t in bct], None, all_Neumann=(1 not in bct))
    P = mg_solver(gpu, m, x)
    rng = np.random.default_rng(11)
    r = np.zeros(g.shape(0), order="F")
    r[1:-1, 1:-1, 1:-1] = rng.standard_normal(tuple(n - 2 for n in g.shape(0)))
    Z = gpu.SF(m, "CC")
    P.apply_prec(gpu.SF(m, "CC").upload(r), Z)
    z = Z.download()
    vol = O.volume(g, 0)
    f = np.zeros(g.shape(0))
    f[1:-1, 1:-1, 1:-1] = r[1:-1, 1:-1, 1:-1] / vol[1:-1, 1:-1, 1:-1]
    zr = MGO.vcycle(levels, f, bct)
    zi, zri = z[1:-1, 1:-1, 1:-1], zr[1:-1, 1:-1, 1:-1]
    err = np.linalg.norm(zi - zri) / np.linalg.norm(zri)
    assert err < 2e-5, err
    assert P.mg_info()[0] == len(levels)


@pytest.mark.parametrize("case", list(CASES))
def test_mixed_pcg_matches_fp64_oracle(gpu, case):
    N, beta, bct = CASES[case]
    hs, g, m = meshes(gpu, N, beta)
    alln = 1 not in bct
    types = [NAMES[t] for t in bct]
    b = np.asfortranarray(O.splitmix(int(np.prod(g.shape(0))), 21).reshape(g.shape(0), order="F"))
    xr = np.zeros(g.shape(0), order="F")
    O.pcg_lap(g, 0, O.BCs.make(types), alln, xr, b, 5000, 1e-13, 1e-300)
    X = gpu.SF(m, "CC").set_BCs(types, None, all_Neumann=alln)
    P = mg_solver(gpu, m, X, tol=1e-10)
    info = P.solve(X, gpu.SF(m, "CC").upload(b))
    assert info.iterations < 40, info.iterations
    assert info.res_norm / info.res_norm0 < 1e-10
    x = X.download()
    vol = O.volume(g, 0)
    xi, xri = x[1:-1, 1:-1, 1:-1], xr[1:-1, 1:-1, 1:-1]
    if alln:
        vi = vol[1:-1, 1:-1, 1:-1]
        xi = xi - (xi * vi).sum() / vi.sum()
        xri = xri - (xri * vi).sum() / vi.sum()
    err = np.linalg.norm(xi - xri) / np.linalg.norm(xri)
    assert err < 1e-6, err


@pytest.mark.parametrize("beta", [None, 1.05])
def test_mixed_pcg_128_reaches_1e8_true_residual(gpu, beta):
    hs, g, m = meshes(gpu, 128, beta)
    types = ["Periodic", "Periodic"] + ["Neumann"] * 4
    b = np.asfortranarray(O.splitmix(int(np.prod(g.shape(0))), 23).reshape(g.shape(0), order="F"))
    X = gpu.SF(m, "CC").set_BCs(types, None, all_Neumann=True)
    P = mg_solver(gpu, m, X, tol=1e-8)
    info = P.solve(X, gpu.SF(m, "CC").upload(b))
    print("128^3 beta", beta, "mixed PCG iterations", info.iterations, flush=True)
    assert info.res_norm / info.res_norm0 < 1e-8
    assert info.iterations < 60
    x = X.download()
    xi = x.copy(order="F")
    O.apply_bcs(g, xi, 0, O.BCs.make(types), True)
    vol = O.volume(g, 0)
    bi, vi = b[1:-1, 1:-1, 1:-1], vol[1:-1, 1:-1, 1:-1]
    mean = (bi * vi).sum() / vi.sum()
    r = (bi - mean - O.lap_centered(g, xi, 0)[1:-1, 1:-1, 1:-1]) * vi
    assert np.linalg.norm(r) / info.res_norm0 < 2e-8
