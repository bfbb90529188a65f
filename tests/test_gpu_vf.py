"""GPU parity of the face-centred Jacobi-PCG (momentum and induction implicit
diffusion) with the oracle: same iteration count, per-solve relative L2 of x
within 1e-10 (the device reductions sum in a different order than the
oracle's sequential sums, so only the dot products differ in rounding), same
exit behaviour at iter_max / tol_abs."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

LDC = [(["Dirichlet"] * 6, [0, 0, 0, 1.0, 0, 0]), (["Dirichlet"] * 6, None), (["Dirichlet"] * 6, None)]
DUCT = [(["Periodic"] * 2 + ["Dirichlet"] * 4, None)] * 3
PSEUDO_VACUUM = [
    (["Neumann", "Neumann"] + ["Dirichlet"] * 4, None),
    (["Dirichlet"] * 2 + ["Neumann", "Neumann"] + ["Dirichlet"] * 2, None),
    (["Dirichlet"] * 4 + ["Neumann", "Neumann"], None),
]


# thin conducting walls (thin_wall, code/handwritten/solvers/induction/init_B_BCs.f90:151-175):
# pseudo-vacuum normal components, Robin tangential ones with c_w = 0.01
CW = [0.01] * 6
THIN_WALL = [
    (["Neumann", "Neumann"] + ["Robin"] * 4, None, CW),
    (["Robin"] * 2 + ["Neumann", "Neumann"] + ["Robin"] * 2, None, CW),
    (["Robin"] * 4 + ["Neumann", "Neumann"], None, CW),
]
# the periodic duct with thin walls (periodic_duct_thin_wall, init_B_BCs.f90:124-134 with x periodic)
DUCT_THIN_WALL = [(["Periodic"] * 2 + t[2:], None, CW) for t, _, _ in THIN_WALL]


def rand(shape, seed, amp=1.0):
    return np.asfortranarray(amp * np.random.default_rng(seed).standard_normal(shape))


def setup(gpu, N, bcspec, seed):
    hs = [O.roberts_both_nodes(0, 1, N[0], 1.1), O.uniform_nodes(0, 2, N[1]), O.roberts_both_nodes(-1, 1, N[2], 1.05)]
    g, m = O.Grid(*hs), gpu.mesh(*hs)
    obcs = [O.BCs.make(*bc) for bc in bcspec]
    x0 = [rand(g.shape(1 << d), seed + d, 0.1) for d in range(3)]
    b = [rand(g.shape(1 << d), seed + 10 + d) for d in range(3)]
    X = [gpu.SF(m, 1 << d).upload(x0[d]).set_BCs(bc[0], bc[1], c_w=bc[2] if len(bc) > 2 else None)
         for d, bc in enumerate(bcspec)]
    B = [gpu.SF(m, 1 << d).upload(b[d]) for d in range(3)]
    return g, m, obcs, x0, b, X, B


def solve_both(gpu, g, m, obcs, x0, b, X, B, op, c, k=None, iter_max=400, tol_rel=1e-10, tol_abs=1e-13, n_skip=1):
    kind = {"mom_diffusion": 1, "ind_diffusion": 2}[op]
    K = None
    if k is not None:
        K = [gpu.SF(m, dl).upload(k[q]) for q, dl in enumerate((6, 5, 3))]
    isp = gpu.iter_solver_params(iter_max=iter_max, tol_rel=tol_rel, tol_abs=tol_abs, n_skip_check_res=n_skip)
    S = gpu.PCG_solver_VF(m, X, isp, op=op, k=K)
    S.update_MFP(c, True)
    info = S.solve(X, B)
    xr = [a.copy(order="F") for a in x0]
    it, res0, res, _ = O.pcg_vf(g, kind, c, obcs, xr, b, iter_max, tol_rel, tol_abs, n_skip, k=k)
    return info, [u.download() for u in X], (it, res0, res, xr)


def rel_l2(a, b):
    num = sum(float(np.sum((x - y) ** 2)) for x, y in zip(a, b))
    den = sum(float(np.sum(y**2)) for y in b)
    return np.sqrt(num / den)


def sigma_inv(g, seed=0):
    rng = np.random.default_rng(seed)
    return [np.asfortranarray(1.0 + 0.5 * rng.random(g.shape(dl))) for dl in (6, 5, 3)]


@pytest.mark.parametrize(
    "op,bcname,N,c",
    [
        ("mom_diffusion", "ldc", (12, 10, 8), -0.02),
        ("mom_diffusion", "duct", (16, 9, 11), -0.05),
        ("ind_diffusion", "pv", (10, 12, 9), 0.03),
        ("ind_diffusion", "duct", (8, 10, 12), 0.01),
        ("ind_diffusion", "thin_wall", (10, 12, 9), 0.03),
        ("ind_diffusion", "duct_thin_wall", (12, 10, 14), 0.02),
    ],
)
def test_vf_pcg_matches_oracle(gpu, op, bcname, N, c):
    spec = {"ldc": LDC, "duct": DUCT, "pv": PSEUDO_VACUUM, "thin_wall": THIN_WALL,
            "duct_thin_wall": DUCT_THIN_WALL}[bcname]
    print(f"vf pcg {op} {bcname} N={N}", flush=True)
    g, m, obcs, x0, b, X, B = setup(gpu, N, spec, 7)
    k = sigma_inv(g) if op == "ind_diffusion" else None
    info, xg, (it, res0, res, xr) = solve_both(gpu, g, m, obcs, x0, b, X, B, op, c, k)
    assert info.iterations == it
    assert abs(info.res_norm0 - res0) <= 1e-12 * res0
    # late residuals sit near the rounding floor of the recurrence: compare on the scale of res0
    assert abs(info.res_norm - res) <= 1e-11 * res0
    assert rel_l2(xg, xr) <= 1e-10


def test_vf_pcg_iter_max_and_n_skip(gpu):
    print("vf pcg iter_max/n_skip", flush=True)
    g, m, obcs, x0, b, X, B = setup(gpu, (10, 9, 8), LDC, 21)
    info, xg, (it, res0, res, xr) = solve_both(gpu, g, m, obcs, x0, b, X, B, "mom_diffusion", -0.3, iter_max=7,
                                                tol_rel=1e-30, n_skip=3)
    assert it == 7 and info.iterations == 7
    assert info.exit_loop & 4
    assert rel_l2(xg, xr) <= 1e-10


def test_vf_pcg_tol_abs_exit(gpu):
    """Zero data and homogeneous BCs: res < tol_abs before the loop, one counted iteration."""
    print("vf pcg tol_abs exit", flush=True)
    N = (6, 6, 6)
    hs = [O.uniform_nodes(0, 1, n) for n in N]
    m = gpu.mesh(*hs)
    X = [gpu.SF(m, 1 << d).fill(0.0).set_BCs(["Dirichlet"] * 6) for d in range(3)]
    B = [gpu.SF(m, 1 << d).fill(0.0) for d in range(3)]
    S = gpu.PCG_solver_VF(m, X, gpu.iter_solver_params(iter_max=50, tol_rel=1e-10, tol_abs=1e-13))
    S.update_MFP(-0.1, True)
    info = S.solve(X, B)
    assert info.iterations == 1 and info.exit_loop & 2
    for u in X:
        assert not np.any(u.download())
