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


EXT_CASES = [(0.3, 5, 1, 1.0), (1.0, 12, 1, 1.0), (1.0, 6, 4, 1.2), (0.05, 4, 1, 1.0)]


def _ext_or_error(fn):
    try:
        return fn()
    except Exception:
        return None


@pytest.mark.parametrize("kind", ["uniform", "near", "far", "B", "prep_uniform", "prep_L", "prep_R", "prep_B",
                                  "prep_C2F", "app_uniform", "app_L", "app_R", "app_B", "app_C2F"])
@pytest.mark.parametrize("L,N,n_iter,ratio", EXT_CASES)
def test_roberts_extensions_match_oracle_and_reference(kind, L, N, n_iter, ratio, tmp_path, monkeypatch):
    """The grid_continue extensions: library == oracle bit for bit, and the
    oracle == the reference's compiled grid_continue bit for bit; where the
    stretch match does not converge all three agree that it fails (the
    reference writes mesh_generation_error.dat, on which MOONS stops at init)."""
    import moons_amd as M
    from oracle import domain_oracle as D

    hn = O.roberts_both_nodes(-1.0, 1.0, 16, 1.05)
    ours = _ext_or_error(lambda: M.nodes_ext(hn, kind, L, N, n_iter, ratio))
    orc = _ext_or_error(lambda: D.ext(hn, kind, L, N, n_iter, ratio))
    assert (ours is None) == (orc is None)
    if ours is not None:
        np.testing.assert_array_equal(ours, orc)
    if O.ref_available():
        monkeypatch.chdir(tmp_path)
        ref = O.ref_ext(hn, kind, L, N, n_iter, ratio)
        if orc is None:
            assert (tmp_path / "mesh_generation_error.dat").exists()
        else:
            assert not (tmp_path / "mesh_generation_error.dat").exists()
            np.testing.assert_array_equal(orc, ref)


def test_roberts_extension_properties():
    """What the stretch matching fixes: the Roberts_B and near extensions
    continue the mesh's end spacing at the junctions; far matches it at the
    outer ends; the original nodes are kept; every spacing is positive; with
    N_iter > 1 a segment that fails the spacing-ratio check gets more cells."""
    import moons_amd as M

    hn = O.uniform_nodes(0.0, 1.0, 10)
    dh = hn[2] - hn[1]
    for kind in ("B", "near"):
        e = M.nodes_ext(hn, kind, 0.5, 3)
        k = 1 + 3  # index of the old hmin
        np.testing.assert_allclose(e[k] - e[k - 1], dh, rtol=1e-12)
        np.testing.assert_allclose(e[-k] - e[-k - 1], dh, rtol=1e-12)
        np.testing.assert_array_equal(e[k:-k], hn[1:-1])
        assert np.all(np.diff(e) > 0)
    f = M.nodes_ext(hn, "far", 1.0, 6)
    np.testing.assert_allclose([f[2] - f[1], f[-2] - f[-3]], dh, rtol=1e-12)
    plain = M.nodes_ext(hn, "B", 1.0, 6)
    auto = M.nodes_ext(hn, "B", 1.0, 6, n_iter=4, max_stretch_ratio=1.2)
    assert auto.size > plain.size
    r = np.diff(auto)
    assert max(np.max(r[1:] / r[:-1]), np.max(r[:-1] / r[1:])) <= 1.2 + 1e-12 or auto.size == plain.size + 2 * 6
    with pytest.raises(M.MoonsError, match="did not converge"):
        M.nodes_ext(hn, "B", 0.05, 4)


def test_generators_reject_bad_arguments():
    import moons_amd as M

    with pytest.raises(M.MoonsError):
        M.nodes_uniform(1.0, 0.0, 8)
    with pytest.raises(M.MoonsError):
        M.nodes_roberts_both(0.0, 1.0, 8, 0.9)
