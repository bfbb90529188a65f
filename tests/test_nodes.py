"""The library's own mesh generators (host code in libmoons_amd.so, no GPU
needed) against the oracle and, where built, the reference's compiled grid
routines (code/handwritten/grid/coordinate_distribution_funcs.f90:192-331,
grid_init.f90:31-68, grid_continue.f90:41-47 and coordinates_extend.f90:529-538
for the ghost nodes).  These generators build every benchmark mesh."""
import os

import numpy as np
import pytest

from oracle import oracle as O

_lib = pytest.importorskip("moons_amd._lib")
pytestmark = pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libmoons_amd.so not built")


@pytest.mark.parametrize("N,lo,hi", [(2, 0.0, 1.0), (7, -1.0, 1.0), (32, 0.0, 1.0), (64, 0.0, 1.0), (128, 0.0, 1.0),
                                     (96, 0.0, 2.0 * np.pi)])
def test_uniform_nodes_match_oracle_and_reference(N, lo, hi):
    import moons_amd as M

    ours = M.nodes_uniform(lo, hi, N)
    np.testing.assert_array_equal(ours, O.uniform_nodes(lo, hi, N))
    if O.ref_available():
        np.testing.assert_array_equal(ours[1:-1], O.ref_grid(0, lo, hi, N))


@pytest.mark.parametrize("N,beta", [(8, 1.05), (32, 1.05), (128, 1.05), (64, 1.2), (16, 3.0), (48, 1.01)])
def test_roberts_nodes_match_oracle_and_reference(N, beta):
    import moons_amd as M

    ours = M.nodes_roberts_both(-1.0, 1.0, N, beta)
    np.testing.assert_array_equal(ours, O.roberts_both_nodes(-1.0, 1.0, N, beta))
    if O.ref_available():
        np.testing.assert_allclose(ours[1:-1], O.ref_grid(1, -1.0, 1.0, N, beta), rtol=0, atol=4e-16)
    # symmetric about the centre, clustered at both walls, ghosts mirror the end cells
    np.testing.assert_allclose(ours + ours[::-1], 0.0, atol=2e-15)
    dh = np.diff(ours[1:-1])
    assert dh[0] < dh[N // 2] and dh[-1] < dh[N // 2]
    np.testing.assert_allclose([ours[1] - ours[0], ours[-1] - ours[-2]], [ours[2] - ours[1], ours[-2] - ours[-3]],
                               rtol=1e-12)


@pytest.mark.parametrize("ext", [1, 3, 11])
def test_extended_nodes_match_reference(ext):
    import moons_amd as M

    hn = O.roberts_both_nodes(0.0, 1.0, 16, 1.1)
    ours = M.nodes_ext_uniform(hn, ext)
    assert ours.size == hn.size + 2 * ext
    if O.ref_available():
        np.testing.assert_array_equal(ours, O.ref_ext_uniform(hn, ext))
    # the fluid nodes are kept, each end's spacing continues
    np.testing.assert_array_equal(ours[ext + 1:-ext - 1], hn[1:-1])
    np.testing.assert_allclose(np.diff(ours[:ext + 2]), hn[2] - hn[1], rtol=1e-12)


def test_generators_reject_bad_arguments():
    import moons_amd as M

    with pytest.raises(M.MoonsError):
        M.nodes_uniform(1.0, 0.0, 8)
    with pytest.raises(M.MoonsError):
        M.nodes_roberts_both(0.0, 1.0, 8, 0.9)
