"""GPU parity of interpolation, edgeCrossFace_E and the constrained-transport
source advect_B against the oracle (bit equality), and div-freeness of the
CT update on the GPU."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(params=["uniform", "roberts"])
def grids(request, gpu):
    if request.param == "uniform":
        hs = [O.uniform_nodes(0, 1, 10)] * 3
    else:
        hs = [O.roberts_both_nodes(0, 1, 9, 1.2), O.uniform_nodes(0, 2, 7), O.roberts_both_nodes(-1, 1, 12, 1.05)]
    return O.Grid(*hs), gpu.mesh(*hs)


def rand(shape, seed):
    return np.asfortranarray(np.random.default_rng(seed).standard_normal(shape))


def test_interp_bitexact(grids, gpu):
    g, m = grids
    seed = 0
    for d in range(3):
        for dl_cc in range(8):
            if (dl_cc >> d) & 1:
                continue
            dl_n = dl_cc | (1 << d)
            for src, dst in ((dl_cc, dl_n), (dl_n, dl_cc)):
                seed += 1
                f = rand(g.shape(src), seed)
                out = gpu.SF(m, dst)
                gpu.interp(out, gpu.SF(m, src).upload(f), d)
                np.testing.assert_array_equal(out.download(), O.interp(g, f, src, dst, d))


def _inputs(g, gpu, m, seed):
    U_E = [rand(g.shape(dl), seed + q) for q, dl in enumerate((6, 6, 5, 5, 3, 3))]
    B = [rand(g.shape(1 << d), seed + 10 + d) for d in range(3)]
    gU = [gpu.SF(m, dl).upload(u) for u, dl in zip(U_E, (6, 6, 5, 5, 3, 3))]
    gB = [gpu.SF(m, 1 << d).upload(B[d]) for d in range(3)]
    return U_E, B, gU, gB


def test_edge_cross_face_E_and_advect_B_bitexact(grids, gpu):
    g, m = grids
    U_E, B, gU, gB = _inputs(g, gpu, m, 50)
    tmp = [gpu.SF(m, dl) for dl in (5, 3, 6, 3, 6, 5)]
    E = [gpu.SF(m, dl) for dl in (6, 5, 3)]
    gpu.edgeCrossFace_E(E, gU, gB, tmp)
    Er = O.edge_cross_face_E(g, U_E, B)
    for d in range(3):
        np.testing.assert_array_equal(E[d].download(), Er[d])
    adv = [gpu.SF(m, 1 << d) for d in range(3)]
    gpu.advect_B(adv, gU, gB, tmp, E)
    ar = O.advect_B(g, U_E, B)
    for d in range(3):
        np.testing.assert_array_equal(adv[d].download(), ar[d])


def test_ct_update_keeps_divB(grids, gpu):
    g, m = grids
    U_E, B, gU, gB = _inputs(g, gpu, m, 70)
    tmp = [gpu.SF(m, dl) for dl in (5, 3, 6, 3, 6, 5)]
    E = [gpu.SF(m, dl) for dl in (6, 5, 3)]
    adv = [gpu.SF(m, 1 << d) for d in range(3)]
    gpu.advect_B(adv, gU, gB, tmp, E)
    dv = gpu.SF(m, "CC")
    gpu.div(dv, *adv)
    a = [x.download() for x in adv]
    assert np.abs(dv.download()[1:-1, 1:-1, 1:-1]).max() < 1e-10 * max(np.abs(x).max() for x in a)
