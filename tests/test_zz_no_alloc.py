"""The allocation-free PPE solve (moons_pcg_sf_solve_noalloc) against the
solver-object solve it mirrors, through the C ABI: two calls in a row, the
second on a smaller mesh (the scratch keeps the first call's contents), each
with the same iteration count, the same residuals and bit-identical x; and
sizes beyond the scratch are refused."""
import ctypes as C

import numpy as np
import pytest

from moons_amd._lib import PcgInfo, load
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def case(gpu, n, beta, seed, types):
    hs = [O.uniform_nodes(0.0, 1.0, n[0]), O.roberts_both_nodes(0.0, 1.0, n[1], beta),
          O.roberts_both_nodes(0.0, 1.0, n[2], beta)]
    m = gpu.mesh(*hs)
    g = O.Grid(*hs)
    b = np.asfortranarray(O.splitmix(int(np.prod(g.shape(0))), seed).reshape(g.shape(0), order="F"))
    alln = all(t in ("Neumann", "Periodic") for t in types)
    return m, b, alln


def solve_object(gpu, m, b, types, alln, isp):
    X = gpu.SF(m, 0).fill(0.0).set_BCs(types, None, all_Neumann=alln)
    P = gpu.PCG_solver_SF(m, X, gpu.iter_solver_params(*isp))
    info = P.solve(X, gpu.SF(m, 0).upload(b))
    return info, X.download()


def solve_noalloc(gpu, m, b, types, alln, isp):
    X = gpu.SF(m, 0).fill(0.0).set_BCs(types, None, all_Neumann=alln)
    B = gpu.SF(m, 0).upload(b)
    info = PcgInfo()
    st = load().moons_pcg_sf_solve_noalloc(X.handle, B.handle, isp[0], isp[1], isp[2], isp[3], C.byref(info))
    assert st == 0, load().moons_last_error().decode()
    return info, X.download()


def test_noalloc_solve_equals_solver_object(gpu):
    isp = (3000, 1e-10, 1e-300, 1)
    cases = [((48, 40, 44), 1.05, 3, ["Periodic", "Periodic"] + ["Neumann"] * 4),
             ((20, 16, 18), 1.1, 4, ["Neumann", "Dirichlet", "Neumann", "Neumann", "Dirichlet", "Neumann"])]
    for n, beta, seed, types in cases:  # the second call is the smaller one
        print(f"noalloc PPE {n} {types}", flush=True)
        m, b, alln = case(gpu, n, beta, seed, types)
        i1, x1 = solve_object(gpu, m, b, types, alln, isp)
        i2, x2 = solve_noalloc(gpu, m, b, types, alln, isp)
        assert i2.iterations == i1.iterations and i2.exit_loop == i1.exit_loop, (i1, i2)
        assert i2.res_norm0 == i1.res_norm0 and i2.res_norm == i1.res_norm
        np.testing.assert_array_equal(x2, x1)


def test_noalloc_refuses_fields_beyond_the_scratch(gpu):
    hs = [O.uniform_nodes(0.0, 1.0, 96)] * 3
    m = gpu.mesh(*hs)
    X = gpu.SF(m, 0).fill(0.0).set_BCs(["Neumann"] * 6, None, all_Neumann=True)
    B = gpu.SF(m, 0).fill(0.0)
    info = PcgInfo()
    st = load().moons_pcg_sf_solve_noalloc(X.handle, B.handle, 10, 1e-8, 1e-300, 1, C.byref(info))
    assert st != 0 and b"scratch" in load().moons_last_error()
