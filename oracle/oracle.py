"""ctypes wrapper of the CPU oracle (liboracle.so) and of the compiled
reference subset (_ref/libmoons_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py -- as the checker, never as the measured path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libmoons_ref.so")

DL = {"CC": 0, "Fx": 1, "Fy": 2, "Fz": 4, "Ex": 6, "Ey": 5, "Ez": 3, "N": 7}
BC = {"none": 0, "Dirichlet": 1, "Neumann": 2, "Periodic": 3, "Symmetric": 4, "Antisymmetric": 5, "Robin": 6}

_dp = C.POINTER(C.c_double)


class BCs(C.Structure):
    _fields_ = [("type", C.c_int * 6), ("val", C.c_double * 6), ("cw", C.c_double * 6)]

    @classmethod
    def make(cls, types, vals=None, c_w=None):
        """Per-face types and values; c_w: the wall conductance ratio of each
        face, used by Robin faces (SP%DP%c_w, init_BC_props)."""
        b = cls()
        for f in range(6):
            t = types[f]
            b.type[f] = BC[t] if isinstance(t, str) else int(t)
            b.val[f] = 0.0 if vals is None else float(vals[f])
            b.cw[f] = 0.0 if c_w is None else float(c_w[f])
        return b


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} missing: run `make -C oracle`")
        L = C.CDLL(ORACLE_SO)
        L.orc_grid_create.restype = C.c_void_p
        L.orc_grid_create.argtypes = [_dp, C.c_int, _dp, C.c_int, _dp, C.c_int]
        L.orc_grid_destroy.argtypes = [C.c_void_p]
        L.orc_field_size.restype = C.c_size_t
        L.orc_field_size.argtypes = [C.c_void_p, C.c_int]
        L.orc_field_shape.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int)]
        L.orc_coord_array.restype = C.c_int
        L.orc_coord_array.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp]
        for nm in ("orc_uniform_nodes",):
            getattr(L, nm).argtypes = [C.c_double, C.c_double, C.c_int, _dp]
        for nm in ("orc_roberts_both_nodes", "orc_roberts_left_nodes", "orc_roberts_right_nodes"):
            getattr(L, nm).argtypes = [C.c_double, C.c_double, C.c_int, C.c_double, _dp]
        L.orc_del.argtypes = [_dp, C.c_int, _dp, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_grad.argtypes = [_dp, _dp, _dp, _dp, C.c_int, C.c_void_p]
        L.orc_div.argtypes = [_dp, C.c_int, _dp, _dp, _dp, C.c_int, C.c_void_p, C.POINTER(BCs)]
        L.orc_curl_component.argtypes = [_dp, C.c_int, _dp, _dp, _dp, C.c_int, C.c_void_p, C.c_int]
        L.orc_lap_centered_sf.argtypes = [_dp, _dp, C.c_int, C.c_void_p, C.POINTER(BCs)]
        L.orc_apply_bcs.argtypes = [_dp, C.c_int, C.c_void_p, C.POINTER(BCs), C.c_int]
        L.orc_grid_face_dh.restype = C.c_double
        L.orc_grid_face_dh.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_volume.argtypes = [_dp, C.c_int, C.c_void_p]
        L.orc_lap_diagonal.argtypes = [_dp, C.c_int, C.c_void_p]
        L.orc_curl_curl_diagonal.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_pcg_lap_sf.restype = C.c_int
        L.orc_pcg_lap_sf.argtypes = [C.c_void_p, C.c_int, C.POINTER(BCs), C.c_int, _dp, _dp, C.c_int, C.c_double,
                                     C.c_double, C.c_int, _dp, _dp]
        L.orc_splitmix_fill.argtypes = [_dp, C.c_size_t, C.c_ulonglong, C.c_double]
        L.orc_nrg_op.argtypes = [_dp, _dp, _dp * 3, C.c_double, C.c_void_p, C.POINTER(BCs)]
        L.orc_pcg_nrg_sf.restype = C.c_int
        L.orc_pcg_nrg_sf.argtypes = [C.c_void_p, C.POINTER(BCs), C.c_int, _dp * 3, C.c_double, _dp, _dp, C.c_int,
                                     C.c_double, C.c_double, C.c_int, _dp, _dp]
        L.orc_interp.argtypes = [_dp, C.c_int, _dp, C.c_int, C.c_void_p, C.c_int]
        L.orc_edge_cross_face_E.argtypes = [_dp, _dp, _dp, _dp * 6, _dp * 3, C.c_void_p]
        L.orc_advect_B.argtypes = [_dp * 3, _dp * 6, _dp * 3, C.c_void_p]
        L.orc_op_vf.argtypes = [_dp * 3, _dp * 3, C.c_int, C.c_double, _dp * 3, C.c_void_p, BCs * 3, C.c_int]
        L.orc_pcg_vf.restype = C.c_int
        L.orc_pcg_vf.argtypes = [C.c_void_p, C.c_int, C.c_double, _dp * 3, BCs * 3, _dp * 3, _dp * 3, C.c_int,
                                 C.c_double, C.c_double, C.c_int, _dp, _dp]
        L.orc_minv_vf.argtypes = [_dp * 3, C.c_int, C.c_double, _dp * 3, C.c_void_p]
        L.orc_assign_ghost_xperiodic.argtypes = [_dp, C.c_int, C.c_void_p, C.POINTER(BCs), C.c_double]
        L.orc_assign_wall_dirichlet.argtypes = [_dp, C.c_int, C.c_void_p, C.POINTER(BCs), C.c_double]
        _lib = L
    return _lib


def p(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def uniform_nodes(hmin, hmax, N):
    hn = np.zeros(N + 3)
    lib().orc_uniform_nodes(hmin, hmax, N, p(hn))
    return hn


def roberts_both_nodes(hmin, hmax, N, beta):
    hn = np.zeros(N + 3)
    lib().orc_roberts_both_nodes(hmin, hmax, N, beta, p(hn))
    return hn


def splitmix(n, seed, amplitude=1.0):
    out = np.zeros(n)
    lib().orc_splitmix_fill(p(out), n, seed, amplitude)
    return out


class Grid:
    """Oracle grid from three node arrays (ghosts included)."""

    def __init__(self, hx, hy, hz):
        self.hn = [np.ascontiguousarray(h, dtype=np.float64) for h in (hx, hy, hz)]
        self.h = lib().orc_grid_create(p(self.hn[0]), len(self.hn[0]), p(self.hn[1]), len(self.hn[1]),
                                       p(self.hn[2]), len(self.hn[2]))
        if not self.h:
            raise ValueError("bad grid")

    def __del__(self):
        try:
            lib().orc_grid_destroy(self.h)
        except Exception:
            pass

    def shape(self, dl):
        dl = DL[dl] if isinstance(dl, str) else dl
        s = (C.c_int * 3)()
        lib().orc_field_shape(self.h, dl, s)
        return (s[0], s[1], s[2])

    def zeros(self, dl):
        return np.zeros(self.shape(dl), order="F")

    def coord(self, d, which):
        n = lib().orc_coord_array(self.h, d, which, None)
        out = np.zeros(n)
        lib().orc_coord_array(self.h, d, which, p(out))
        return out


def face_dh(g: Grid, d: int, is_max: int) -> float:
    """dh of the face sub-domain that the Neumann and Robin rules use."""
    return lib().orc_grid_face_dh(g.h, d, is_max)


def _F(a):
    assert a.flags.f_contiguous and a.dtype == np.float64
    return p(a)


def del_op(g: Grid, out, dl_out, f, dl_in, n, d, pad, gen):
    lib().orc_del(_F(out), dl_out, _F(f), dl_in, g.h, n, d, pad, gen)


def grad(g: Grid, u, dl_u):
    outs = [g.zeros(dl_u ^ (1 << d)) for d in range(3)]
    lib().orc_grad(_F(outs[0]), _F(outs[1]), _F(outs[2]), _F(u), dl_u, g.h)
    return outs


def div(g: Grid, u, v, w, dl_out, bcs=None):
    out = g.zeros(dl_out)
    lib().orc_div(_F(out), dl_out, _F(u), _F(v), _F(w), dl_out ^ 1, g.h, C.byref(bcs) if bcs else None)
    return out


def curl(g: Grid, u, v, w, dl_out, d):
    out = g.zeros(dl_out)
    lib().orc_curl_component(_F(out), dl_out, _F(u), _F(v), _F(w), 0, g.h, d)
    return out


def lap_centered(g: Grid, U, dl, bcs=None):
    out = g.zeros(dl)
    lib().orc_lap_centered_sf(_F(out), _F(U), dl, g.h, C.byref(bcs) if bcs else None)
    return out


def interp(g: Grid, f, dl_in, dl_out, d):
    out = g.zeros(dl_out)
    lib().orc_interp(_F(out), dl_out, _F(f), dl_in, g.h, d)
    return out


def edge_cross_face_E(g: Grid, U_E, B):
    E = [g.zeros(dl) for dl in (6, 5, 3)]
    lib().orc_edge_cross_face_E(_F(E[0]), _F(E[1]), _F(E[2]), (_dp * 6)(*[_F(u) for u in U_E]),
                                (_dp * 3)(*[_F(b) for b in B]), g.h)
    return E


def advect_B(g: Grid, U_E, B):
    adv = [g.zeros(1 << d) for d in range(3)]
    lib().orc_advect_B((_dp * 3)(*[_F(a) for a in adv]), (_dp * 6)(*[_F(u) for u in U_E]),
                       (_dp * 3)(*[_F(b) for b in B]), g.h)
    return adv


MOM_DIFFUSION, IND_DIFFUSION = 1, 2


def _p3(arrs):
    return (_dp * 3)(*[_F(a) for a in arrs])


def _k3(g, k):
    if k is None:  # unused by mom_diffusion; any edge arrays do
        k = [g.zeros(dl) for dl in (6, 5, 3)]
    return k, _p3(k)


def op_vf(g: Grid, x, kind, c, bcs3, k=None, implicit=True):
    """Ax of mom_diffusion / ind_diffusion; x is modified (implicit BCs) when implicit."""
    Ax = [g.zeros(1 << d) for d in range(3)]
    k, kp = _k3(g, k)
    lib().orc_op_vf(_p3(Ax), _p3(x), kind, c, kp, g.h, (BCs * 3)(*bcs3), 1 if implicit else 0)
    return Ax


def minv_vf(g: Grid, kind, c, k=None):
    M = [g.zeros(1 << d) for d in range(3)]
    k, kp = _k3(g, k)
    lib().orc_minv_vf(_p3(M), kind, c, kp, g.h)
    return M


def pcg_vf(g: Grid, kind, c, bcs3, x, b, iter_max, tol_rel, tol_abs, n_skip=1, k=None):
    """VF PCG; x (three F-ordered face arrays) updated in place."""
    out = np.zeros(4)
    hist = np.zeros(iter_max + 2)
    k, kp = _k3(g, k)
    st = lib().orc_pcg_vf(g.h, kind, c, kp, (BCs * 3)(*bcs3), _p3(x), _p3(b), iter_max, tol_rel, tol_abs,
                          n_skip, p(out), p(hist))
    if st != 0:
        raise RuntimeError("oracle pcg_vf failed")
    return int(out[0]), out[1], out[2], hist[: int(out[3])].copy()


def nrg_op(g: Grid, x, k, c, bcs: BCs):
    """Ax of nrg_diffusion (c div(k grad x) + x); x gets its implicit BCs."""
    Ax = g.zeros(0)
    lib().orc_nrg_op(_F(Ax), _F(x), _p3(k), c, g.h, C.byref(bcs))
    return Ax


def pcg_nrg(g: Grid, bcs: BCs, all_neumann, k, c, x, b, iter_max, tol_rel, tol_abs, n_skip=1):
    """Energy PCG (nrg_diffusion, prec_lap_SF); x updated in place."""
    out = np.zeros(4)
    hist = np.zeros(iter_max + 2)
    st = lib().orc_pcg_nrg_sf(g.h, C.byref(bcs), 1 if all_neumann else 0, _p3(k), c, _F(x), _F(b), iter_max,
                              tol_rel, tol_abs, n_skip, p(out), p(hist))
    if st != 0:
        raise RuntimeError("oracle pcg_nrg failed")
    return int(out[0]), out[1], out[2], hist[: int(out[3])].copy()


def apply_bcs(g: Grid, u, dl, bcs: BCs, implicit=False):
    lib().orc_apply_bcs(_F(u), dl, g.h, C.byref(bcs), 1 if implicit else 0)


def volume(g: Grid, dl):
    out = g.zeros(dl)
    lib().orc_volume(_F(out), dl, g.h)
    return out


def curl_curl_diagonal(g: Grid, sig):
    """The oracle's curl-curl diagonal: sig on the x, y, z edges -> x, y, z faces."""
    sig = [np.asfortranarray(a, dtype=np.float64) for a in sig]
    outs = [np.zeros(g.shape(1 << d), order="F") for d in range(3)]
    lib().orc_curl_curl_diagonal(_p3(outs), _p3(sig), g.h)
    return outs


def lap_diagonal(g: Grid, dl):
    out = g.zeros(dl)
    lib().orc_lap_diagonal(_F(out), dl, g.h)
    return out


def set_threads(n: int) -> None:
    """Threads of the C oracle's OpenMP loops.  1 (the default) keeps every
    sum in its sequential order; more threads split the stencils and the
    element-wise loops (unchanged results) and add padded sums per plane (last
    bits differ), for the CPU timing baseline."""
    lib().orc_set_threads(int(n))


def pcg_lap(g: Grid, dl, bcs: BCs, all_neumann, x, b, iter_max, tol_rel, tol_abs, n_skip=1):
    """Returns (iterations, res0, res, res_history); x is updated in place."""
    out = np.zeros(4)
    hist = np.zeros(iter_max + 2)
    st = lib().orc_pcg_lap_sf(g.h, dl, C.byref(bcs), 1 if all_neumann else 0, _F(x), _F(b), iter_max, tol_rel,
                              tol_abs, n_skip, p(out), p(hist))
    if st != 0:
        raise RuntimeError("oracle pcg failed")
    return int(out[0]), out[1], out[2], hist[: int(out[3])].copy()


# ------------------------------------------------------------- compiled reference
_ref = None


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    """The reference's own routines (flang-built from its sources), lazily bound."""
    global _ref
    if _ref is None:
        R = C.CDLL(REF_SO, mode=os.RTLD_LAZY)
        _ip = C.POINTER(C.c_int)
        R.ref_grid.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, C.c_double, _dp]
        R.ref_stencil.argtypes = [C.c_int, _dp, C.c_int, _dp, _dp, _dp, _ip, _ip, _ip]
        R.ref_interp_stencil.argtypes = [_dp, C.c_int, _dp, C.c_int, _dp, _dp, _ip]
        R.ref_pencil.argtypes = [C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, _dp, _dp, C.c_int, C.c_int]
        R.ref_bc_plane.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_double, C.c_int, C.c_int]
        R.ref_gf_export.argtypes = [_dp, _ip, C.c_char_p, C.c_int]
        R.ref_gf_import.argtypes = [C.c_char_p, C.c_int, _dp, C.c_int, _ip]
        raw = [_dp, C.c_int, _dp, C.c_int, _dp, C.c_int, C.c_int, _dp, C.c_int, C.c_char_p, C.c_int,
               C.c_char_p, C.c_int, C.c_int]
        R.ref_export_raw.argtypes = raw
        R.ref_import_raw.argtypes = raw
        R.ref_ext_uniform.argtypes = [_dp, C.c_int, C.c_int, _dp]
        R.ref_ext.argtypes = [C.c_int, _dp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_double, _dp, C.c_int,
                              C.POINTER(C.c_int)]
        R.ref_overlap.argtypes = [C.c_int, _dp, C.c_int, _dp, C.c_int, C.c_int, _ip]
        R.ref_robin_plane.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_int,
                                      C.c_int]
        R.ref_face_dh.argtypes = [_dp, C.c_int, C.c_int, _dp]
        grid3 = [_dp, C.c_int, _dp, C.c_int, _dp, C.c_int]
        R.ref_laplacian_diagonal.argtypes = grid3 + [C.c_int, _dp, C.c_int]
        R.ref_curl_curl_diagonal.argtypes = grid3 + [_dp, C.c_int] * 6
        R.ref_interp.argtypes = grid3 + [C.c_int, C.c_int, C.c_int, _dp, C.c_int, _dp, C.c_int]
        _ref = R
    return _ref


def ref_grid(kind, hmin, hmax, N, beta=0.0):
    hn = np.zeros(N + 1)
    ref().ref_grid(kind, hmin, hmax, N, beta, p(hn))
    return hn


def ref_stencil(kind, dh, s):
    dh = np.ascontiguousarray(dh, dtype=np.float64)
    L, D, U = np.zeros(s), np.zeros(s), np.zeros(s)
    nL, nD, nU = C.c_int(), C.c_int(), C.c_int()
    ref().ref_stencil(kind, p(dh), s, p(L), p(D), p(U), C.byref(nL), C.byref(nD), C.byref(nU))
    return L[: nL.value], D[: nD.value], U[: nU.value]


def ref_interp(hc, hn):
    hc = np.ascontiguousarray(hc, dtype=np.float64)
    hn = np.ascontiguousarray(hn, dtype=np.float64)
    D, U = np.zeros(len(hc)), np.zeros(len(hc))
    n = C.c_int()
    ref().ref_interp_stencil(p(hc), len(hc), p(hn), len(hn), p(D), p(U), C.byref(n))
    return D[: n.value], U[: n.value]


def ref_pencil(kind, dfdh, f, L, D, U, gt):
    dfdh = np.ascontiguousarray(dfdh, dtype=np.float64).copy()
    f = np.ascontiguousarray(f, dtype=np.float64)
    nc = len(D)
    Lp = np.ascontiguousarray(L if len(L) == nc else np.zeros(nc), dtype=np.float64)
    ref().ref_pencil(kind, p(dfdh), len(dfdh), p(f), len(f), p(Lp), p(np.ascontiguousarray(D)),
                     p(np.ascontiguousarray(U)), nc, gt)
    return dfdh


def ref_bc_plane(kind, ug, ub, ui, bvals, dh_nhat):
    ug = np.array(ug, dtype=np.float64, order="F")
    ub = np.array(ub, dtype=np.float64, order="F")
    ui = np.asfortranarray(ui, dtype=np.float64)
    bvals = np.asfortranarray(bvals, dtype=np.float64)
    x, y = ug.shape
    ref().ref_bc_plane(kind, p(ug), p(ub), p(ui), p(bvals), dh_nhat, x, y)
    return ug, ub


def ref_robin_plane(kind, ug, ub, ui, bvals, dh, nhat, c_w):
    """The reference's raw Robin kernels (kind 0 Robin_C, 1 Robin_N,
    2 Robin_C_implicit, 3 Robin_N_implicit) on one ghost plane."""
    ug = np.array(ug, dtype=np.float64, order="F")
    ub = np.asfortranarray(ub, dtype=np.float64)
    ui = np.asfortranarray(ui, dtype=np.float64)
    bvals = np.asfortranarray(bvals, dtype=np.float64)
    x, y = ug.shape
    ref().ref_robin_plane(kind, p(ug), p(ub), p(ui), p(bvals), float(dh), float(nhat), float(c_w), x, y)
    return ug


def ref_face_dh(hn, face: int) -> float:
    """FSD%dh of face 1..6 on a grid with node array hn in every direction."""
    hn = np.ascontiguousarray(hn, dtype=np.float64)
    out = C.c_double(0.0)
    ref().ref_face_dh(p(hn), len(hn), int(face), C.byref(out))
    return out.value


def _grid_args(hs):
    hs = [np.ascontiguousarray(h, dtype=np.float64) for h in hs]
    return hs, [a for h in hs for a in (p(h), len(h))]


def ref_laplacian_diagonal(g: Grid, dl: int) -> np.ndarray:
    """laplacian_diagonal of the reference on g at data location dl."""
    hs, ga = _grid_args(g.hn)
    out = np.zeros(g.shape(dl), order="F")
    ref().ref_laplacian_diagonal(*ga, dl, _F(out), out.size)
    return out


def ref_curl_curl_diagonal(g: Grid, sig):
    """curl_curl_diagonal of the reference: sig on the x, y, z edges -> x, y, z faces."""
    hs, ga = _grid_args(g.hn)
    sig = [np.asfortranarray(a, dtype=np.float64) for a in sig]
    outs = [np.zeros(g.shape(1 << d), order="F") for d in range(3)]
    args = []
    for a in sig + outs:
        args += [_F(a), a.size]
    ref().ref_curl_curl_diagonal(*ga, *args)
    return outs


def ref_interp_field(g: Grid, f, dl_in: int, dl_out: int, d: int) -> np.ndarray:
    """interpO2_SF's per-block interpolation (C2N with end-plane extrapolation,
    or N2C) of the reference along d (0-based)."""
    hs, ga = _grid_args(g.hn)
    f = np.asfortranarray(f, dtype=np.float64)
    out = np.zeros(g.shape(dl_out), order="F")
    ref().ref_interp(*ga, dl_out, dl_in, d + 1, _F(f), f.size, _F(out), out.size)
    return out


def ref_gf_export(path: str, f: np.ndarray) -> None:
    """The reference's grid_field export of the 3D array f (column-major)."""
    f = np.asfortranarray(f, dtype=np.float64)
    s = (C.c_int * 3)(*f.shape)
    b = path.encode()
    ref().ref_gf_export(f.ctypes.data_as(_dp), s, b, len(b))


def ref_gf_import(path: str, n: int):
    """The reference's grid_field import: (shape, flat column-major values)."""
    f = np.zeros(max(n, 1))
    s = (C.c_int * 3)()
    b = path.encode()
    ref().ref_gf_import(b, len(b), p(f), n, s)
    return tuple(s), f[:n]


def _raw_args(hn, dl, f, directory, name, pad):
    hx, hy, hz = (np.ascontiguousarray(h, dtype=np.float64) for h in hn)
    fl = np.asfortranarray(f, dtype=np.float64).ravel(order="F").copy()
    d, nm = directory.encode(), name.encode()
    return [p(hx), len(hx), p(hy), len(hy), p(hz), len(hz), dl, p(fl), fl.size, d, len(d), nm, len(nm), pad], fl


def ref_export_raw(hn, dl, f, directory, name, pad) -> None:
    """The reference's steady raw Tecplot export of f on the grid of node arrays hn."""
    args, _ = _raw_args(hn, dl, f, directory, name, pad)
    ref().ref_export_raw(*args)


def ref_import_raw(hn, dl, f, directory, name, pad) -> np.ndarray:
    """The reference's raw Tecplot import into a copy of f (pad cells kept)."""
    args, fl = _raw_args(hn, dl, f, directory, name, pad)
    ref().ref_import_raw(*args)
    return fl.reshape(np.shape(f), order="F")


REF_EXT_KINDS = ("uniform", "near", "far", "B", "prep_uniform", "prep_L", "prep_R", "prep_B", "prep_C2F",
                 "app_uniform", "app_L", "app_R", "app_B", "app_C2F")


def ref_ext(hn, kind: str, L: float, N: int, n_iter: int = 1, max_ratio: float = 1.0) -> np.ndarray:
    """The reference's grid_continue extension `kind` of the node array hn
    (ghosts included).  A stretch match that does not converge makes the
    reference write the file mesh_generation_error in the working directory."""
    hn = np.ascontiguousarray(hn, dtype=np.float64)
    cap = hn.size + 8 * (N + n_iter + 4)
    out = np.zeros(cap)
    n = C.c_int(0)
    ref().ref_ext(REF_EXT_KINDS.index(kind), p(hn), hn.size, float(L), int(N), int(n_iter), float(max_ratio), p(out),
                  cap, C.byref(n))
    assert n.value <= cap
    return out[: n.value].copy()


def ref_ext_uniform(hn, N: int) -> np.ndarray:
    """The reference's ext_uniform_IO of the node array hn (with ghosts)."""
    hn = np.ascontiguousarray(hn, dtype=np.float64)
    out = np.zeros(hn.size + 2 * N)
    ref().ref_ext_uniform(p(hn), hn.size, N, p(out))
    return out


def ref_overlap(kind: str, hn1, hn2, pad: int = 1):
    """The reference's get_C_overlap / get_N_overlap: (i1 lo, i1 hi, i2 lo, i2 hi), 1-based."""
    a = np.ascontiguousarray(hn1, dtype=np.float64)
    b = np.ascontiguousarray(hn2, dtype=np.float64)
    r = (C.c_int * 4)()
    ref().ref_overlap(0 if kind == "C" else 1, p(a), a.size, p(b), b.size, pad, r)
    return tuple(r)
