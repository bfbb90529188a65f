"""GPU parity of the staggered operators and BC kernels (through the C ABI)
against the oracle on the same seeded inputs.  Element-wise kernels are
compiled without FP contraction and follow the reference's operation order,
so the bar is bit equality."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

GRIDS = {
    "uniform8": [("u", 8), ("u", 8), ("u", 8)],
    "roberts_aniso": [("r", 9, 1.2), ("u", 6), ("r", 11, 1.05)],
    "uniform_16x12x10": [("u", 16), ("u", 12), ("u", 10)],
}


def nodes(spec):
    if spec[0] == "u":
        return O.uniform_nodes(0.0, 1.0, spec[1])
    return O.roberts_both_nodes(0.0, 1.0, spec[1], spec[2])


@pytest.fixture(params=list(GRIDS))
def grids(request, gpu):
    hs = [nodes(s) for s in GRIDS[request.param]]
    return O.Grid(*hs), gpu.mesh(*hs)


def rand(shape, seed):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.standard_normal(shape))


def test_mesh_coordinates_match(grids, gpu):
    g, m = grids
    import ctypes as C
    for d in range(3):
        for which in range(4):
            n = C.c_int()
            gpu.load().moons_mesh_coords(m.handle, d, which, None, C.byref(n))
            out = np.zeros(n.value)
            gpu.load().moons_mesh_coords(m.handle, d, which, out.ctypes.data_as(C.POINTER(C.c_double)), None)
            np.testing.assert_array_equal(out, g.coord(d, which))


@pytest.mark.parametrize("gen", [1, 2, 3])
@pytest.mark.parametrize("pad", [0, 1])
def test_del_all_kinds_bitexact(grids, gpu, gen, pad):
    g, m = grids
    seed = 100
    for d in range(3):
        for dl_in, dl_out, n in ((0, 1 << d, 1), (1 << d, 0, 1), (0, 0, 1), (0, 0, 2), (7, 7, 1), (7, 7 ^ (1 << d), 1),
                                 (6, 6, 2)):
            if (dl_in & ~(1 << d)) != (dl_out & ~(1 << d)):
                continue
            seed += 1
            f = rand(g.shape(dl_in), seed)
            out0 = rand(g.shape(dl_out), seed + 1000)
            ref = out0.copy(order="F")
            O.del_op(g, ref, dl_out, f, dl_in, n, d, pad, gen)
            F = gpu.SF(m, dl_in).upload(f)
            R = gpu.SF(m, dl_out).upload(out0)
            gpu.del_op(R, F, n, d, pad, gen)
            np.testing.assert_array_equal(R.download(), ref, err_msg=f"d={d} {dl_in}->{dl_out} n={n}")


def test_grad_div_curl_bitexact(grids, gpu):
    g, m = grids
    u = rand(g.shape(0), 1)
    U = gpu.SF(m, 0).upload(u)
    G = [gpu.SF(m, 1 << d) for d in range(3)]
    gpu.grad(*G, U)
    ref = O.grad(g, u, 0)
    for d in range(3):
        np.testing.assert_array_equal(G[d].download(), ref[d])
    fx = [rand(g.shape(1 << d), 10 + d) for d in range(3)]
    Fx = [gpu.SF(m, 1 << d).upload(fx[d]) for d in range(3)]
    D = gpu.SF(m, 0).fill(7.0)
    gpu.div(D, *Fx)
    np.testing.assert_array_equal(D.download(), O.div(g, fx[0], fx[1], fx[2], 0))
    # curl of a face field lands on edges, of an edge field on faces
    for src, dsts in (((1, 2, 4), (6, 5, 3)), ((6, 5, 3), (1, 2, 4))):
        v = [rand(g.shape(dl), 20 + dl) for dl in src]
        V = [gpu.SF(m, dl).upload(a) for dl, a in zip(src, v)]
        for d, dl_out in enumerate(dsts):
            C_ = gpu.SF(m, dl_out)
            gpu.curl(C_, *V, d)
            np.testing.assert_array_equal(C_.download(), O.curl(g, v[0], v[1], v[2], dl_out, d))


BC_SETS = [
    (["Neumann"] * 6, [0.0] * 6),
    (["Dirichlet"] * 6, [0.3, -0.2, 1.0, 0.0, 0.5, -1.5]),
    (["Periodic", "Periodic", "Neumann", "Dirichlet", "Neumann", "Neumann"], [0, 0, 0.2, 0.7, -0.1, 0.4]),
    (["Symmetric", "Antisymmetric", "Dirichlet", "Neumann", "Periodic", "Periodic"], [0, 0, 1.0, 0.25, 0, 0]),
]


@pytest.mark.parametrize("bc", range(len(BC_SETS)))
@pytest.mark.parametrize("implicit", [False, True])
def test_apply_bcs_bitexact(grids, gpu, bc, implicit):
    g, m = grids
    types, vals = BC_SETS[bc]
    for dl in (0, 1, 2, 4, 3, 5, 6, 7):
        u = rand(g.shape(dl), 40 + dl)
        ref = u.copy(order="F")
        O.apply_bcs(g, ref, dl, O.BCs.make(types, vals), implicit)
        U = gpu.SF(m, dl).upload(u).set_BCs(types, vals)
        (gpu.apply_BCs_implicit if implicit else gpu.apply_BCs)(U)
        np.testing.assert_array_equal(U.download(), ref, err_msg=f"dl={dl}")


ROBIN_SETS = [
    (["Robin"] * 6, [0.3, -0.2, 1.0, 0.0, 0.5, -1.5], [0.01, 0.02, 0.5, 1.0, 3.0, 0.07]),
    (["Periodic", "Periodic", "Robin", "Neumann", "Dirichlet", "Robin"], [0, 0, 0.2, 0.7, -0.1, 0.4],
     [0, 0, 0.01, 0, 0, 0.01]),
]


@pytest.mark.parametrize("bc", range(len(ROBIN_SETS)))
@pytest.mark.parametrize("implicit", [False, True])
def test_apply_robin_bcs_bitexact(grids, gpu, bc, implicit):
    """Thin-wall (Robin) ghost rules with each face's c_w, every location."""
    g, m = grids
    types, vals, cw = ROBIN_SETS[bc]
    for dl in (0, 1, 2, 4, 3, 5, 6, 7):
        u = rand(g.shape(dl), 60 + dl)
        ref = u.copy(order="F")
        O.apply_bcs(g, ref, dl, O.BCs.make(types, vals, cw), implicit)
        U = gpu.SF(m, dl).upload(u).set_BCs(types, vals, c_w=cw)
        (gpu.apply_BCs_implicit if implicit else gpu.apply_BCs)(U)
        np.testing.assert_array_equal(U.download(), ref, err_msg=f"dl={dl}")
        assert not np.array_equal(ref, u)


@pytest.mark.parametrize("bc", [0, 1, 2])
def test_lap_uniform_sf_bitexact(grids, gpu, bc):
    """Lap_uniform_SF: implicit BCs on x, then div(grad x) -- fused kernel for CC."""
    g, m = grids
    types, vals = BC_SETS[bc]
    x = rand(g.shape(0), 77)
    xr = x.copy(order="F")
    O.apply_bcs(g, xr, 0, O.BCs.make(types, vals), True)
    ref = O.lap_centered(g, xr, 0)
    X = gpu.SF(m, 0).upload(x).set_BCs(types, vals)
    A = gpu.SF(m, 0).fill(3.0)
    gpu.Lap_uniform_SF(A, X)
    np.testing.assert_array_equal(A.download(), ref)
    np.testing.assert_array_equal(X.download(), xr)  # x's ghosts hold the implicit BC values


def test_volume_bitexact(grids, gpu):
    g, m = grids
    for dl in range(8):
        V = gpu.SF(m, dl)
        gpu.volume(V)
        np.testing.assert_array_equal(V.download(), O.volume(g, dl))
