"""The Fortran host surface (moons_amd/fortran, flang-built) drives the GPU
pressure-Poisson solve through the C ABI; its result matches the oracle."""
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
    out = subprocess.run([DRIVER, str(N), str(beta), str(prec), "3000", repr(tol), str(bpath), str(xpath)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "PCG iterations" in out.stdout
    x = np.fromfile(xpath, dtype=np.float64).reshape(shape, order="F")
    xr = np.zeros(shape, order="F")
    O.pcg_lap(g, 0, O.BCs.make(["Neumann"] * 6), True, xr, b, 3000, 1e-10 if prec == 1 else 1e-13, 1e-300)
    vol = O.volume(g, 0)[1:-1, 1:-1, 1:-1]
    xi, xri = x[1:-1, 1:-1, 1:-1], xr[1:-1, 1:-1, 1:-1]
    if prec == 2:
        xi = xi - (xi * vol).sum() / vol.sum()
        xri = xri - (xri * vol).sum() / vol.sum()
    err = np.linalg.norm(xi - xri) / np.linalg.norm(xri)
    assert err < bar, err
