"""GPU parity of the cosine-transform direct solver and clean_div_FFT with
the restatement (oracle/fft_oracle.py: transforms as cosine matrices, pinned
against the reference's FFT-based DCT; the reference's mode factors by
default, the exact eigenvalues as the opt-in), and at 128^3 the size-independent
property: the solution satisfies the 5-point Neumann Laplacian of every plane
with right-hand side f minus its plane mean.  The GPU transforms are FP64
GEMMs, the restatement sums in another order, so fields are compared within
1e-12 relative L2."""
import numpy as np
import pytest

from oracle import fft_oracle as FO
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def grid(n, lengths):
    return [O.uniform_nodes(0.0, L, N) for N, L in zip(n, lengths)]


BCS = {
    "neumann": (["Neumann"] * 6, None),
    "mixed": (["Neumann", "Dirichlet", "Neumann", "Neumann", "Dirichlet", "Neumann"], [0.0, 0.3, 0.0, 0.0, -0.2, 0.1]),
}


def first_spacings(hs):
    return [float(hn[1] - hn[0]) for hn in hs]


@pytest.mark.parametrize("coefficients", ["reference", "exact"])
@pytest.mark.parametrize("direction", [1, 2, 3])
@pytest.mark.parametrize("bc", list(BCS))
def test_fft_solve_matches_restatement(gpu, direction, bc, coefficients):
    """Unequal transverse spacings, where the reference's mode factors and the
    exact eigenvalues differ; both modes against their restatement."""
    print(f"fft solve dir {direction} {bc} {coefficients}", flush=True)
    hs = grid((16, 12, 8), (1.0, 0.75, 1.25))
    g = O.Grid(*hs)
    m = gpu.mesh(*hs)
    h = [float(np.diff(hn)[1]) for hn in hs]
    types, vals = BCS[bc]
    f = np.asfortranarray(O.splitmix(int(np.prod(g.shape(0))), 11 + direction).reshape(g.shape(0), order="F"))
    U = gpu.SF(m, 0).fill(0.0).set_BCs(types, vals, all_Neumann=bc == "neumann")
    F = gpu.SF(m, 0).upload(f)
    S = gpu.FFT_solver_SF(m, direction, coefficients=coefficients)
    S.solve(U, F)
    ref = np.zeros(g.shape(0), order="F")
    h0 = first_spacings(hs) if coefficients == "reference" else None
    ref[1:-1, 1:-1, 1:-1] = FO.solve_interior(f[1:-1, 1:-1, 1:-1], h, direction, h_first=h0)
    O.apply_bcs(g, ref, 0, O.BCs.make(types, vals))
    assert rel_l2(U.download(), ref) < 1e-12


def test_fft_rejects_stretched_transverse_grid(gpu):
    hs = [O.uniform_nodes(0.0, 1.0, 8), O.roberts_both_nodes(0.0, 1.0, 8, 1.1), O.uniform_nodes(0.0, 1.0, 8)]
    m = gpu.mesh(*hs)
    with pytest.raises(gpu.MoonsError):
        gpu.FFT_solver_SF(m, 1)
    gpu.FFT_solver_SF(m, 2)  # y is the solve direction: x and z are uniform


def test_clean_div_fft_matches_restatement(gpu):
    print("clean_div_FFT", flush=True)
    hs = grid((12, 16, 10), (1.0, 1.0, 0.8))
    g = O.Grid(*hs)
    m = gpu.mesh(*hs)
    h = [float(np.diff(hn)[1]) for hn in hs]
    rng = np.random.default_rng(5)
    us = [np.asfortranarray(rng.standard_normal(g.shape(1 << d)) * 1e-3) for d in range(3)]
    ubc = (["Dirichlet"] * 6, [0.0, 0.0, 0.0, 1.0, 0.0, 0.0])
    ptypes = ["Neumann"] * 6
    scale = 1.0 / 1e-3
    direction = 3
    # restated sequence (clean_divergence.f90:61-83), the reference's mode factors
    dv = O.div(g, us[0], us[1], us[2], 0) * scale
    phi_r = np.zeros(g.shape(0), order="F")
    phi_r[1:-1, 1:-1, 1:-1] = FO.solve_interior(dv[1:-1, 1:-1, 1:-1], h, direction, h_first=first_spacings(hs))
    O.apply_bcs(g, phi_r, 0, O.BCs.make(ptypes))
    gr = O.grad(g, phi_r, 0)
    Xr = []
    for d in range(3):
        x = np.asfortranarray(us[d] - gr[d] * (1.0 / scale))
        O.apply_bcs(g, x, 1 << d, O.BCs.make(*ubc))
        Xr.append(x)
    Xs = [gpu.SF(m, 1 << d).upload(us[d]) for d in range(3)]
    X = [gpu.SF(m, 1 << d).set_BCs(*ubc) for d in range(3)]
    T = [gpu.SF(m, 1 << d) for d in range(3)]
    phi = gpu.SF(m, 0).set_BCs(ptypes, None, all_Neumann=True)
    tcc = gpu.SF(m, 0)
    S = gpu.FFT_solver_SF(m, direction)
    gpu.clean_div_FFT(S, X, Xs, phi, scale, T, tcc)
    assert rel_l2(phi.download(), phi_r) < 1e-12
    for d in range(3):
        assert rel_l2(X[d].download(), Xr[d]) < 1e-12


@pytest.mark.parametrize("coefficients", ["reference", "exact"])
def test_fft_128_plane_laplacian(gpu, coefficients):
    """128^3, direction 1, equal transverse spacings: the GPU solution
    satisfies the Neumann Laplacian of every x-plane (the oracle's staggered
    derivatives) with f minus its plane mean, and every plane of u has zero
    mean -- with the reference's factors on every plane but the first, whose
    modes (0, q) the reference's guard leaves untouched."""
    print(f"fft 128 {coefficients}", flush=True)
    N = 128
    hs = grid((N, N, N), (2.0, 1.0, 1.0))
    g = O.Grid(*hs)
    m = gpu.mesh(*hs)
    f = np.asfortranarray(O.splitmix(int(np.prod(g.shape(0))), 21).reshape(g.shape(0), order="F"))
    U = gpu.SF(m, 0).fill(0.0).set_BCs(["Neumann"] * 6, None, all_Neumann=True)
    S = gpu.FFT_solver_SF(m, 1, coefficients=coefficients)
    S.solve(U, gpu.SF(m, 0).upload(f))
    u = np.asfortranarray(U.download())
    out = g.zeros(0)
    for q, e in enumerate((1, 2)):
        t = g.zeros(1 << e)
        O.del_op(g, t, 1 << e, u, 0, 1, e, 0, 1)
        O.del_op(g, out, 0, t, 1 << e, 1, e, 0, 1 if q == 0 else 2)
    fi = f[1:-1, 1:-1, 1:-1]
    expect = fi - fi.mean(axis=(1, 2), keepdims=True)
    first = 1 if coefficients == "reference" else 0
    assert np.abs(out[1:-1, 1:-1, 1:-1][first:] - expect[first:]).max() < 1e-9 * np.abs(fi).max()
    assert np.abs(u[1:-1, 1:-1, 1:-1].mean(axis=(1, 2))).max() < 1e-12 * np.abs(u).max()
