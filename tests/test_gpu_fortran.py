"""The Fortran host surface (moons_amd/fortran, flang-built) drives the GPU
pressure-Poisson solve, the face-centred implicit diffusion solves and the
energy step through the C ABI; each result matches the oracle."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "moons_amd", "fortran", "_build", "ppe_driver")


@pytest.mark.parametrize("prec,tol,bar", [(1, 1e-10, 1e-10), (2, 1e-10, 1e-6)])
def test_fortran_driver_ppe(tmp_path, prec, tol, bar):
    if not os.path.exists(DRIVER):
        pytest.fail("Fortran driver not built (run __graft_entry__.build())")
    N, beta = 24, 1.05
    hx = O.uniform_nodes(0.0, 1.0, N)
    hy = O.roberts_both_nodes(0.0, 1.0, N, beta)
    g = O.Grid(hx, hy, hy)
    shape = g.shape(0)
    b = np.asfortranarray(O.splitmix(int(np.prod(shape)), 31).reshape(shape, order="F"))
    bpath, xpath = tmp_path / "b.bin", tmp_path / "x.bin"
    b.ravel(order="F").tofile(bpath)
    out = subprocess.run([DRIVER, str(N), str(beta), str(prec), "3000", repr(tol), str(bpath), str(xpath),
                          str(tmp_path) + "/"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "PCG iterations" in out.stdout
    x = np.fromfile(xpath, dtype=np.float64).reshape(shape, order="F")
    # export_raw / export of the device solution from Fortran read back bit for bit
    import moons_amd as M
    xb = x.view(np.int64)
    np.testing.assert_array_equal(M.raw_read(np.zeros(shape), "CC", str(tmp_path) + "/", "x", 0).view(np.int64), xb)
    np.testing.assert_array_equal(M.gf_read(str(tmp_path / "x_restart.txt")).view(np.int64), xb)
    xr = np.zeros(shape, order="F")
    O.pcg_lap(g, 0, O.BCs.make(["Neumann"] * 6), True, xr, b, 3000, 1e-10 if prec == 1 else 1e-13, 1e-300)
    vol = O.volume(g, 0)[1:-1, 1:-1, 1:-1]
    xi, xri = x[1:-1, 1:-1, 1:-1], xr[1:-1, 1:-1, 1:-1]
    if prec == 2:
        xi = xi - (xi * vol).sum() / vol.sum()
        xri = xri - (xri * vol).sum() / vol.sum()
    err = np.linalg.norm(xi - xri) / np.linalg.norm(xri)
    assert err < bar, err


def test_fortran_driver_fft(tmp_path):
    """init_FFT / solve from Fortran host arrays (ppe_driver prec 0): the x-plane
    cosine-transform solve, with the reference's mode factors, equals the
    restatement."""
    from oracle import fft_oracle as FO

    if not os.path.exists(DRIVER):
        pytest.fail("Fortran driver not built (run __graft_entry__.build())")
    N = 16
    hx = O.uniform_nodes(0.0, 1.0, N)
    g = O.Grid(hx, hx, hx)
    shape = g.shape(0)
    b = np.asfortranarray(O.splitmix(int(np.prod(shape)), 32).reshape(shape, order="F"))
    bpath, xpath = tmp_path / "b.bin", tmp_path / "x.bin"
    b.ravel(order="F").tofile(bpath)
    out = subprocess.run([DRIVER, str(N), "0", "0", "1", "1e-8", str(bpath), str(xpath)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "FFT solve done" in out.stdout
    x = np.fromfile(xpath, dtype=np.float64).reshape(shape, order="F")
    h = float(np.diff(hx)[1])
    xr = np.zeros(shape, order="F")
    h0 = float(hx[1] - hx[0])
    xr[1:-1, 1:-1, 1:-1] = FO.solve_interior(b[1:-1, 1:-1, 1:-1], [h, h, h], 1, h_first=[h0, h0, h0])
    O.apply_bcs(g, xr, 0, O.BCs.make(["Neumann"] * 6))
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-12


VF_DRIVER = os.path.join(ROOT, "moons_amd", "fortran", "_build", "vf_driver")


@pytest.mark.parametrize("op,c", [(1, -0.02), (2, 0.03)])
def test_fortran_driver_vf(tmp_path, op, c):
    """init_PCG_VF / update_MFP / solve from Fortran host arrays: momentum with
    the lid-driven-cavity BCs, induction with pseudo-vacuum BCs and sigmaInv."""
    if not os.path.exists(VF_DRIVER):
        pytest.fail("Fortran VF driver not built (run __graft_entry__.build())")
    N, beta = 12, 1.1
    hx = O.uniform_nodes(0.0, 1.0, N)
    hy = O.roberts_both_nodes(0.0, 1.0, N, beta)
    g = O.Grid(hx, hy, hy)
    rng = np.random.default_rng(op)
    x0 = [np.asfortranarray(0.1 * rng.standard_normal(g.shape(1 << d))) for d in range(3)]
    b = [np.asfortranarray(rng.standard_normal(g.shape(1 << d))) for d in range(3)]
    paths = {k: tmp_path / f"{k}.bin" for k in ("x0", "b", "x", "sig")}
    np.concatenate([a.ravel(order="F") for a in x0]).tofile(paths["x0"])
    np.concatenate([a.ravel(order="F") for a in b]).tofile(paths["b"])
    if op == 1:
        bcs = [O.BCs.make(["Dirichlet"] * 6, [0, 0, 0, 1.0, 0, 0]), O.BCs.make(["Dirichlet"] * 6),
               O.BCs.make(["Dirichlet"] * 6)]
        k = None
    else:
        bcs = [O.BCs.make(["Neumann", "Neumann"] + ["Dirichlet"] * 4),
               O.BCs.make(["Dirichlet"] * 2 + ["Neumann", "Neumann"] + ["Dirichlet"] * 2),
               O.BCs.make(["Dirichlet"] * 4 + ["Neumann", "Neumann"])]
        k = [np.asfortranarray(1.0 + 0.5 * rng.random(g.shape(dl))) for dl in (6, 5, 3)]
        np.concatenate([a.ravel(order="F") for a in k]).tofile(paths["sig"])
    args = [VF_DRIVER, str(N), str(beta), str(op), repr(c), "400", "1e-10", str(paths["x0"]), str(paths["b"]),
            str(paths["x"])] + ([str(paths["sig"])] if op == 2 else [])
    out = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "PCG_VF iterations" in out.stdout
    flat = np.fromfile(paths["x"], dtype=np.float64)
    xs, off = [], 0
    for d in range(3):
        n = int(np.prod(g.shape(1 << d)))
        xs.append(flat[off:off + n].reshape(g.shape(1 << d), order="F"))
        off += n
    xr = [a.copy(order="F") for a in x0]
    O.pcg_vf(g, op, c, bcs, xr, b, 400, 1e-10, 1e-13, 1, k=k)
    num = sum(float(np.sum((p - q) ** 2)) for p, q in zip(xs, xr))
    den = sum(float(np.sum(q ** 2)) for q in xr)
    assert np.sqrt(num / den) <= 1e-10


NRG_DRIVER = os.path.join(ROOT, "moons_amd", "fortran", "_build", "nrg_driver")


@pytest.mark.parametrize("method,rk", [(3, 1), (5, 3)])
def test_fortran_driver_energy(tmp_path, method, rk):
    """init(nrg) / set_joule / solve(nrg, T, U) from Fortran host arrays, every
    energy source on, 3 steps: T within 1e-10 of the oracle's step."""
    from oracle import step_oracle as S

    if not os.path.exists(NRG_DRIVER):
        pytest.fail("Fortran driver not built (run __graft_entry__.build())")
    N, beta = 12, 1.05
    hx = O.uniform_nodes(0.0, 1.0, N)
    hy = O.roberts_both_nodes(0.0, 1.0, N, beta)
    g = O.Grid(hx, hy, hy)
    rng = np.random.default_rng(11)
    tb = O.BCs.make(["Dirichlet", "Dirichlet"] + ["Neumann"] * 4, [1.0, 0.0, 0.0, 0.0, 0.0, 0.0])
    X = np.meshgrid(g.coord(0, 1), g.coord(1, 1), g.coord(2, 1), indexing="ij")[0]
    T0 = np.asfortranarray(1.0 - X + 0.05 * np.sin(7.0 * X))
    O.apply_bcs(g, T0, 0, tb)
    U = [np.asfortranarray(0.3 * (2 * rng.random(g.shape(1 << d)) - 1)) for d in range(3)]
    J = [np.asfortranarray(rng.standard_normal(g.shape(S.EDGE_DL[d]))) for d in range(3)]
    sig = np.asfortranarray(0.5 + rng.random(g.shape(0)))
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    np.concatenate([a.ravel(order="F") for a in [T0] + U + J + [sig]]).tofile(inp)
    out = subprocess.run([NRG_DRIVER, str(N), str(beta), str(method), str(rk), "3", str(inp), str(outp)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "energy T-solve iterations" in out.stdout
    res = np.fromfile(outp, dtype=np.float64)
    n = int(np.prod(g.shape(0)))
    T, F = res[:n].reshape(g.shape(0), order="F"), res[n:].reshape(g.shape(0), order="F")
    k = [np.full(g.shape(1 << d), (1.0, 1.5, 2.0)[d], order="F") for d in range(3)]
    ost = S.EnergyState(g, T0, tb, k, dt=2e-3, method=method, coeff_implicit=-1.0 / 50.0, advection=(True, -1.0),
                        diffusion=(True, 1.0 / 50.0), isp=(50, 1e-12, 1e-15, 1), KE_diffusion=(True, 0.02),
                        viscous_dissipation=(True, 0.5), joule_heating=(True, 0.25), volumetric_heating=(True, 0.7),
                        rk_stages=rk, J=J, sigma_inv_cc=sig)
    for _ in range(3):
        ost.step(U)
    assert np.linalg.norm(T - ost.T) / np.linalg.norm(ost.T) <= 1e-10
    assert np.linalg.norm(F - ost.F) / np.linalg.norm(ost.F) <= 1e-10
