"""Python host mirror of the MOONS SF / mesh / PCG call surface over libmoons_amd.

Names follow the reference (code/handwritten): `SF` (scalar field,
fields/SF_extend.f90), `mesh` built from per-direction node coordinates
(grid/coordinates_extend.f90:78-173), `PCG_solver_SF` with init/solve
(PCG/PCG_solver_extend.f90:51-206), `clean_div` (solvers/clean_divergence.f90:37-59),
and the discrete operators grad/div/curl/lap (ops/ops_discrete.f90).  All
compute runs in the HIP library on the GPU; this module only holds handles.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import MhdInfo, MhdParams, NrgInfo, NrgParams, PcgInfo, check, load

DL = {"CC": 0, "Fx": 1, "Fy": 2, "Fz": 4, "Ex": 6, "Ey": 5, "Ez": 3, "N": 7}
BC = {"none": 0, "Dirichlet": 1, "Neumann": 2, "Periodic": 3, "Symmetric": 4, "Antisymmetric": 5, "Robin": 6}

_initialised = False


def init(device: int = 0) -> None:
    """Select the GPU and create the library stream (once per process)."""
    global _initialised
    check(load().moons_init(int(device)))
    _initialised = True


def synchronize() -> None:
    check(load().moons_synchronize())


KTIMERS = {"mg_line_fine": 1, "pcg_matvec": 2, "pcg_update": 3, "pcg_pupdate": 4, "unet_conv16": 5}


class ktimer:
    """In-situ kernel timer: `with ktimer("mg_line_fine") as t: ...` times
    every launch of that kernel class on the library stream inside the block;
    afterwards t.launches and t.avg_ms hold the count and mean duration."""

    def __init__(self, which: str, max_launches: int = 1 << 16):
        self.which, self.max = KTIMERS[which], int(max_launches)
        self.launches, self.avg_ms = 0, 0.0

    def __enter__(self):
        check(load().moons_ktimer_start(self.which, self.max))
        return self

    def __exit__(self, *exc):
        n, a = C.c_int(0), C.c_double(0.0)
        check(load().moons_ktimer_stop(C.byref(n), C.byref(a)))
        self.launches, self.avg_ms = n.value, a.value
        return False


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def nodes_uniform(hmin: float, hmax: float, N: int) -> np.ndarray:
    """uniform_func + ghost nodes (N+3 node coordinates)."""
    hn = np.zeros(N + 3)
    check(load().moons_nodes_uniform(hmin, hmax, N, _dp(hn)))
    return hn


def nodes_roberts_both(hmin: float, hmax: float, N: int, beta: float) -> np.ndarray:
    """robertsBoth_func (Tannehill transformation 2, alpha=1/2) + ghost nodes."""
    hn = np.zeros(N + 3)
    check(load().moons_nodes_roberts_both(hmin, hmax, N, beta, _dp(hn)))
    return hn


def nodes_ext_uniform(hn, N: int) -> np.ndarray:
    """ext_uniform_IO: the node array (with ghosts) extended by N uniform
    cells at both ends (code/handwritten/grid/grid_continue.f90:41-47)."""
    hn = np.ascontiguousarray(hn, dtype=np.float64)
    out = np.zeros(hn.size + 2 * N)
    check(load().moons_nodes_ext_uniform(_dp(hn), hn.size, N, _dp(out)))
    return out


EXT_KINDS = ("uniform", "near", "far", "B", "prep_uniform", "prep_L", "prep_R", "prep_B", "prep_C2F",
             "app_uniform", "app_L", "app_R", "app_B", "app_C2F")


def nodes_ext(hn, kind: str, L: float = 0.0, N: int = 1, n_iter: int = 1, max_stretch_ratio: float = 1.0) -> np.ndarray:
    """A grid_continue extension of the node array hn (ghosts included), by
    kind: "uniform" (ext_uniform_IO), "near" / "far" / "B"
    (ext_Roberts_near/far/B_IO), "prep_uniform", "prep_L", "prep_R", "prep_B",
    "prep_C2F" (ext_prep_*_IO) and "app_*" likewise
    (code/handwritten/grid/grid_continue.f90): segments of N cells and length L
    whose stretching is matched to the mesh's end spacing; n_iter and
    max_stretch_ratio are the mesh-quality parameters N_iter and
    max_mesh_stretch_ratio."""
    hn = np.ascontiguousarray(hn, dtype=np.float64)
    cap = hn.size + 2 * (N + n_iter + 2)
    out = np.zeros(cap)
    n = C.c_int(0)
    check(load().moons_nodes_ext(_dp(hn), hn.size, EXT_KINDS.index(kind), float(L), int(N), int(n_iter),
                                 float(max_stretch_ratio), _dp(out), cap, C.byref(n)))
    return out[: n.value].copy()


class mesh:
    """Single-block mesh from three node-coordinate arrays (ghosts included)."""

    def __init__(self, hn_x, hn_y, hn_z):
        if not _initialised:
            init(0)
        self.hn = [np.ascontiguousarray(h, dtype=np.float64) for h in (hn_x, hn_y, hn_z)]
        h = C.c_void_p()
        check(load().moons_mesh_create(_dp(self.hn[0]), len(self.hn[0]), _dp(self.hn[1]), len(self.hn[1]),
                                       _dp(self.hn[2]), len(self.hn[2]), C.byref(h)))
        self.handle = h

    @classmethod
    def cube_uniform(cls, N: int, hmin: float = 0.0, hmax: float = 1.0) -> "mesh":
        hn = nodes_uniform(hmin, hmax, N)
        return cls(hn, hn, hn)

    @property
    def sn(self):
        return [len(h) for h in self.hn]

    def shape(self, dl: int):
        return tuple(self.sn[d] if (dl >> d) & 1 else self.sn[d] - 1 for d in range(3))

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                load().moons_mesh_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class SF:
    """Device-resident scalar field of one block (reference SF, single BF)."""

    def __init__(self, m: mesh, dl="CC"):
        self.m = m
        self.dl = DL[dl] if isinstance(dl, str) else int(dl)
        h = C.c_void_p()
        check(load().moons_field_create(m.handle, self.dl, C.byref(h)))
        self.handle = h
        s = (C.c_int * 3)()
        check(load().moons_field_shape(h, s))
        self.shape = (s[0], s[1], s[2])

    @property
    def size(self) -> int:
        return int(np.prod(self.shape))

    def set_BCs(self, types, vals=None, all_Neumann: bool = False, c_w=None) -> "SF":
        """Per-face BC types and values; c_w: the wall conductance ratio of
        each face, which Robin faces use (init_BC_props)."""
        t = (C.c_int * 6)(*[BC[x] if isinstance(x, str) else int(x) for x in types])
        v = (C.c_double * 6)(*(vals if vals is not None else [0.0] * 6))
        check(load().moons_field_set_bcs(self.handle, t, v, 1 if all_Neumann else 0))
        if c_w is not None:
            self.set_BC_props(c_w)
        return self

    def set_BC_props(self, c_w, Robin_coeff=None) -> "SF":
        """init_BC_props: c_w per face, and the Robin coefficients kept beside it."""
        cw = (C.c_double * 6)(*[float(x) for x in c_w])
        rc = (C.c_double * 6)(*[float(x) for x in (Robin_coeff if Robin_coeff is not None else [0.0] * 6)])
        check(load().moons_field_set_bc_props(self.handle, cw, rc))
        return self

    def upload(self, a: np.ndarray) -> "SF":
        a = np.asfortranarray(a, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError(f"shape {a.shape} != field shape {self.shape}")
        flat = a.ravel(order="F")
        check(load().moons_field_upload(self.handle, _dp(flat), flat.size))
        return self

    def download(self) -> np.ndarray:
        flat = np.empty(self.size, dtype=np.float64)
        check(load().moons_field_download(self.handle, _dp(flat), flat.size))
        return flat.reshape(self.shape, order="F")

    def fill(self, v: float) -> "SF":
        check(load().moons_field_fill(self.handle, float(v)))
        return self

    def assign(self, other: "SF") -> "SF":
        check(load().moons_field_copy(self.handle, other.handle))
        return self

    def export(self, path: str) -> None:
        """Writes the field's grid_field restart primitives
        (export_primitives_grid_field, code/generated_code/grid_field/grid_field.f90:123-139)."""
        check(load().moons_field_export(self.handle, os.fsencode(path)))

    def import_(self, path: str) -> "SF":
        """Reads grid_field restart primitives of this field's shape
        (import_primitives_grid_field, code/generated_code/grid_field/grid_field.f90:141-156)."""
        check(load().moons_field_import(self.handle, os.fsencode(path)))
        return self

    @classmethod
    def borrow(cls, m: mesh, handle) -> "SF":
        """View of a field owned by the library (not destroyed by this object)."""
        f = cls.__new__(cls)
        f.m = m
        f.handle = C.c_void_p(handle)
        f._borrowed = True
        s = (C.c_int * 3)()
        check(load().moons_field_shape(f.handle, s))
        f.shape = (s[0], s[1], s[2])
        f.dl = -1
        return f

    def __del__(self):
        try:
            if getattr(self, "handle", None) and not getattr(self, "_borrowed", False):
                load().moons_field_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


# ------------------------------------------------------ multi-domain meshes
class mesh_domain:
    """Where the interior mesh lies in the total mesh (init_MD,
    code/handwritten/mesh_domain/mesh_domain_extend.f90:31-38)."""

    def __init__(self, m_interior: "mesh", m_total: "mesh"):
        h = C.c_void_p()
        check(load().moons_domain_create(m_interior.handle, m_total.handle, C.byref(h)))
        self.handle = h
        self._keep = (m_interior, m_total)

    def overlap(self, dir: int, node: bool):
        """(interior lo, interior hi, total lo, total hi), 0-based inclusive."""
        r = (C.c_int * 4)()
        check(load().moons_domain_overlap(self.handle, dir, 1 if node else 0, r))
        return tuple(r)

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                load().moons_domain_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def embed(total, interior, MD: mesh_domain) -> None:
    """embed (code/handwritten/ops/ops_embedExtract.f90:169-278): an SF or lists of SFs."""
    for t, i in zip(_raw_fields(total), _raw_fields(interior)):
        check(load().moons_embed(t.handle, i.handle, MD.handle))


def extract(interior, total, MD: mesh_domain) -> None:
    """extract (code/handwritten/ops/ops_embedExtract.f90:49-165): an SF or lists of SFs."""
    for i, t in zip(_raw_fields(interior), _raw_fields(total)):
        check(load().moons_extract(i.handle, t.handle, MD.handle))


def set_sigma_inv(sigma_inv, MD_sigma: mesh_domain, sig_local_over_sig_f: float) -> None:
    """set_sigma_inv_SF/VF (code/handwritten/solvers/induction/induction_aux.f90:162-194)."""
    for f in _raw_fields(sigma_inv):
        check(load().moons_set_sigma_inv(f.handle, MD_sigma.handle, sig_local_over_sig_f))


# ------------------------------------------ restart primitives, raw export
def gf_write(path: str, f: np.ndarray) -> None:
    """grid_field restart primitives of a host array (column-major order)."""
    f = np.asfortranarray(f, dtype=np.float64)
    s = (C.c_int * 3)(*f.shape)
    flat = f.ravel(order="F")
    check(load().moons_gf_write(os.fsencode(path), _dp(flat), s))


def gf_read(path: str) -> np.ndarray:
    """The array of a grid_field restart-primitives file (shape from the file)."""
    s = (C.c_int * 3)()
    check(load().moons_gf_read(os.fsencode(path), None, 0, s))
    shape = (s[0], s[1], s[2])
    if min(shape) <= 0:
        return np.zeros((0, 0, 0))
    flat = np.empty(int(np.prod(shape)))
    check(load().moons_gf_read(os.fsencode(path), _dp(flat), flat.size, s))
    return flat.reshape(shape, order="F")


def _raw_fields(x):
    return list(x) if isinstance(x, (list, tuple)) else [x]


def export_raw(m: mesh, x, dir: str, name: str, pad: int, t: float | None = None) -> None:
    """Tecplot raw export of an SF, or of each component of a staggered VF
    given as three SFs (er_steady_SF / er_steady_VF / the unsteady forms with
    t, code/handwritten/IO/export_raw_processed.f90:44-84)."""
    tt = C.c_double(t) if t is not None else None
    for f in _raw_fields(x):
        check(load().moons_export_raw(f.handle, os.fsencode(dir), name.encode(), int(pad),
                                      C.byref(tt) if tt is not None else None))


def import_raw(m: mesh, x, dir: str, name: str, pad: int) -> None:
    """Raw Tecplot import into an SF or each of three SFs (ir_steady_SF /
    ir_steady_VF, code/handwritten/IO/import_raw.f90:23-43)."""
    for f in _raw_fields(x):
        check(load().moons_import_raw(f.handle, os.fsencode(dir), name.encode(), int(pad)))


def raw_write(hn, dl, f: np.ndarray, dir: str, name: str, pad: int, t: float | None = None) -> None:
    """export_raw of a host array f on the grid of node arrays hn (with ghosts)."""
    hx, hy, hz = (np.ascontiguousarray(h, dtype=np.float64) for h in hn)
    flat = np.asfortranarray(f, dtype=np.float64).ravel(order="F")
    d = DL[dl] if isinstance(dl, str) else int(dl)
    tt = C.c_double(t) if t is not None else None
    check(load().moons_raw_write(_dp(hx), hx.size, _dp(hy), hy.size, _dp(hz), hz.size, d, _dp(flat),
                                 os.fsencode(dir), name.encode(), int(pad), C.byref(tt) if tt is not None else None))


def raw_read(f: np.ndarray, dl, dir: str, name: str, pad: int) -> np.ndarray:
    """import_raw into a copy of the host array f (its pad cells kept)."""
    a = np.array(f, dtype=np.float64, order="F")
    s = (C.c_int * 3)(*a.shape)
    flat = a.ravel(order="F")
    d = DL[dl] if isinstance(dl, str) else int(dl)
    check(load().moons_raw_read(s, d, _dp(flat), os.fsencode(dir), name.encode(), int(pad)))
    return flat.reshape(a.shape, order="F")


# ------------------------------------------------------------------ operators
def apply_BCs(u: SF) -> None:
    check(load().moons_apply_bcs(u.handle))


def apply_BCs_implicit(u: SF) -> None:
    check(load().moons_apply_bcs_implicit(u.handle))


def del_op(dfdh: SF, f: SF, n: int, dir: int, pad: int, gen: int) -> None:
    """del%assign (gen=1) / add (2) / subtract (3), dir 0..2 (ops_del.f90)."""
    check(load().moons_del(dfdh.handle, f.handle, n, dir, pad, gen))


def grad(gx: SF, gy: SF, gz: SF, u: SF) -> None:
    check(load().moons_grad(gx.handle, gy.handle, gz.handle, u.handle))


def div(divU: SF, u: SF, v: SF, w: SF) -> None:
    check(load().moons_div(divU.handle, u.handle, v.handle, w.handle))


def curl(curlU: SF, u: SF, v: SF, w: SF, dir: int) -> None:
    check(load().moons_curl(curlU.handle, u.handle, v.handle, w.handle, dir))


def Lap_uniform_SF(Ax: SF, x: SF) -> None:
    check(load().moons_lap_uniform_sf(Ax.handle, x.handle))


def interp(out: SF, f: SF, dir: int) -> None:
    """interpO2_SF: C->N (theta weights, extrapolated ends) or N->C along dir."""
    check(load().moons_interp(out.handle, f.handle, dir))


def _arr(fields):
    return (C.c_void_p * len(fields))(*[f.handle for f in fields])


def edgeCrossFace_E(E, U_E, B, tmp) -> None:
    """E = U_E x face2Edge_no_diag(B) (ops_advect.f90:202-218)."""
    check(load().moons_edge_cross_face_E(_arr(E), _arr(U_E), _arr(B), _arr(tmp)))


def advect_B(adv, U_E, B, temp_E_TF, temp_E) -> None:
    """adv = curl(U_E x B) on faces (ops_advect.f90:186-200)."""
    check(load().moons_advect_B(_arr(adv), _arr(U_E), _arr(B), _arr(temp_E_TF), _arr(temp_E)))


def volume(vol: SF) -> None:
    check(load().moons_volume(vol.handle))


# ------------------------------------------------------------------ solvers
@dataclass
class iter_solver_params:
    """ISP: exit criteria of an iterative solve (var_set/iter_solver_params_extend.f90)."""

    iter_max: int = 1000
    tol_rel: float = 1e-6
    tol_abs: float = 1e-13
    n_skip_check_res: int = 1


@dataclass
class pcg_info:
    iterations: int
    exit_loop: int
    res_norm0: float
    res_norm: float
    ms: float

    @property
    def rel(self) -> float:
        return self.res_norm / self.res_norm0 if self.res_norm0 else 0.0


class PCG_solver_SF:
    """PCG for a cell-centred field.

    op="Lap_uniform_SF": the Laplacian (pressure Poisson, phi cleaning).
    op="nrg_diffusion": I + c div(k grad) of the energy equation; set k with
    set_k and c with update_MFP; prec="Lap_SF" only.
    prec="Lap_SF": the reference's Jacobi preconditioner (prec_Lap_SF), FP64.
    prec="MG_FP32": one FP32 multigrid V-cycle inside an FP64 flexible PCG.
    """

    OPS = {"Lap_uniform_SF": 1, "nrg_diffusion": 3}
    PRECS = {"Lap_SF": 1, "MG_FP32": 2, "UNet": 3}

    def __init__(self, m: mesh, x: SF, ISP: iter_solver_params | None = None, prec: str = "Lap_SF",
                 op: str = "Lap_uniform_SF"):
        h = C.c_void_p()
        check(load().moons_pcg_sf_create(m.handle, self.OPS[op], self.PRECS[prec], x.handle, C.byref(h)))
        self.handle = h
        self.m = m
        self.prec = prec
        self.set_ISP(ISP or iter_solver_params())

    def set_k(self, k) -> None:
        """The nrg_diffusion operator's face coefficients (three face SFs, copied)."""
        check(load().moons_pcg_sf_set_k(self.handle, _arr(k)))

    def update_MFP(self, coeff_implicit_time_split: float, update_preconditioner: bool = True) -> None:
        check(load().moons_pcg_sf_update_mfp(self.handle, coeff_implicit_time_split, 1 if update_preconditioner else 0))

    def set_MG(self, max_levels: int = 16, nu1: int = 2, nu2: int = 2, coarse_sweeps: int = 16) -> None:
        check(load().moons_pcg_set_mg(self.handle, max_levels, nu1, nu2, coarse_sweeps))

    def mg_info(self):
        lv = C.c_int()
        n = (C.c_int * 3)()
        check(load().moons_pcg_mg_info(self.handle, C.byref(lv), n))
        return lv.value, (n[0], n[1], n[2])

    def set_MG_UNet(self, net: "UNet | None", level: int = 2, cycles: int = 0) -> None:
        """Learned initial guess of multigrid level `level` inside the V-cycle
        (prec="MG_FP32"), in the first `cycles` V-cycles of each solve (0: every one)."""
        check(load().moons_pcg_set_mg_unet(self.handle, net.handle if net is not None else None, level))
        check(load().moons_pcg_set_mg_unet_cycles(self.handle, int(cycles)))
        self._mg_net = net

    def set_MG_gamma(self, level: int, cycles: int) -> None:
        """Cycle the V-cycle's part from multigrid level `level` down `cycles`
        times per visit (prec="MG_FP32"; level < 0 or cycles = 1: plain V-cycle)."""
        check(load().moons_pcg_set_mg_gamma(self.handle, int(level), int(cycles)))

    def set_fallback(self, enable: bool) -> None:
        """Jacobi redo of failed learned solves (on by default)."""
        check(load().moons_pcg_set_fallback(self.handle, 1 if enable else 0))

    def set_UNet(self, net: "UNet") -> None:
        """Bind the learned preconditioner (prec="UNet")."""
        check(load().moons_pcg_set_unet(self.handle, net.handle))
        self._net = net

    def apply_prec(self, r: SF, z: SF) -> None:
        check(load().moons_pcg_apply_prec(self.handle, r.handle, z.handle))

    def set_ISP(self, ISP: iter_solver_params) -> None:
        self.ISP = ISP
        check(load().moons_pcg_set_isp(self.handle, ISP.iter_max, ISP.tol_rel, ISP.tol_abs, ISP.n_skip_check_res))

    def solve(self, x: SF, b: SF) -> pcg_info:
        info = PcgInfo()
        check(load().moons_pcg_sf_solve(self.handle, x.handle, b.handle, C.byref(info)))
        return pcg_info(info.iterations, info.exit_loop, info.res_norm0, info.res_norm, info.ms)

    def time_kernels(self, reps: int = 20):
        ms = (C.c_double * 3)()
        check(load().moons_pcg_time_kernels(self.handle, reps, ms))
        return [ms[0], ms[1], ms[2]]

    def time_MG(self, reps: int = 20):
        """ms per launch of the finest half-sweep, restriction, prolongation,
        coarse end, a whole V-cycle, the finest fused sweep, and the half-sweep
        and fused sweep of levels 1 and 2 (multigrid solvers only)."""
        ms = (C.c_double * 10)()
        check(load().moons_pcg_time_mg(self.handle, reps, ms))
        return list(ms)

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                load().moons_pcg_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def clean_div(PCG: PCG_solver_SF, X, Xstar, phi: SF, scale: float, temp_F, temp_CC: SF) -> pcg_info:
    """X = Xstar - grad(phi)/scale with lap(phi) = scale*div(Xstar) (clean_divergence.f90:37-59)."""
    arr = C.c_void_p * 3
    info = PcgInfo()
    check(load().moons_clean_div(PCG.handle, arr(*[u.handle for u in X]), arr(*[u.handle for u in Xstar]), phi.handle,
                                 float(scale), arr(*[u.handle for u in temp_F]), temp_CC.handle, C.byref(info)))
    return pcg_info(info.iterations, info.exit_loop, info.res_norm0, info.res_norm, info.ms)


class FFT_solver_SF:
    """The cosine-transform direct solver (init_FFT / solve,
    code/handwritten/FFT_solver/FFT_solver_extend.f90:43-137): every plane
    normal to `direction` (1, 2, 3) is an independent 2D Poisson problem solved
    by DCT-II transforms; the transverse spacings must be uniform.
    coefficients: "reference" (init_FFT's mode factors, the default) or
    "exact" (the 5-point Laplacian's eigenvalues for any spacings)."""

    COEFS = {"reference": 0, "exact": 1}

    def __init__(self, m: mesh, direction: int, coefficients: str = "reference"):
        h = C.c_void_p()
        check(load().moons_fft_create(m.handle, int(direction), C.byref(h)))
        self.handle = h
        self.m = m
        if coefficients != "reference":
            check(load().moons_fft_set_coefficients(h, self.COEFS[coefficients]))

    def solve(self, u: SF, f: SF) -> None:
        check(load().moons_fft_solve(self.handle, u.handle, f.handle))

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                load().moons_fft_destroy(self.handle)
        except Exception:
            pass


def clean_div_FFT(FFT: FFT_solver_SF, X, Xstar, phi: SF, scale: float, temp_F, temp_CC: SF) -> None:
    """clean_div_FFT (code/handwritten/solvers/clean_divergence.f90:61-83)."""
    check(load().moons_clean_div_fft(FFT.handle, _arr(X), _arr(Xstar), phi.handle, scale, _arr(temp_F),
                                     temp_CC.handle))


class PCG_solver_VF:
    """Jacobi-PCG for the face-centred implicit diffusion solves (init_PCG_VF,
    code/handwritten/PCG/PCG_solver_extend.f90:117-180).

    op="mom_diffusion": A = I + c lap (prec_mom_VF); op="ind_diffusion":
    A = I + c curl(k curl) with k = sigmaInv on edges (prec_ind_VF).  x is the
    three face components whose BCs the solver keeps.  Call update_MFP before
    solving (the coefficient starts at 0).
    """

    OPS = {"mom_diffusion": 1, "ind_diffusion": 2}

    def __init__(self, m: mesh, x, ISP: iter_solver_params | None = None, op: str = "mom_diffusion", k=None):
        h = C.c_void_p()
        karg = _arr(k) if k is not None else None
        check(load().moons_pcg_vf_create(m.handle, self.OPS[op], _arr(x), karg, C.byref(h)))
        self.handle = h
        self.m = m
        self.op = op
        self.set_ISP(ISP or iter_solver_params())

    def set_ISP(self, ISP: iter_solver_params) -> None:
        self.ISP = ISP
        check(load().moons_pcg_vf_set_isp(self.handle, ISP.iter_max, ISP.tol_rel, ISP.tol_abs, ISP.n_skip_check_res))

    def update_MFP(self, coeff_implicit_time_split: float, update_preconditioner: bool = True) -> None:
        check(load().moons_pcg_vf_update_mfp(self.handle, float(coeff_implicit_time_split),
                                             1 if update_preconditioner else 0))

    def solve(self, x, b) -> pcg_info:
        info = PcgInfo()
        check(load().moons_pcg_vf_solve(self.handle, _arr(x), _arr(b), C.byref(info)))
        return pcg_info(info.iterations, info.exit_loop, info.res_norm0, info.res_norm, info.ms)

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                load().moons_pcg_vf_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def uniform_gravity_field(m: mesh, direction: int):
    """The reference's uniform_gravity_field (code/handwritten/solvers/energy/
    init_gravity_field.f90:41-50): three cell-centred fields, component
    direction (1..3) equal to 1 everywhere, ghosts included, the others 0."""
    if direction not in (1, 2, 3):
        raise ValueError("uniform_gravity_field: direction must be 1, 2 or 3")
    return [SF(m, "CC").fill(1.0 if d == direction - 1 else 0.0) for d in range(3)]


class MHD_step:
    """One MOONS time step on the device (solve_MOONS,
    code/handwritten/MOONS/MOONS_solver.f90:47-144): momentum sources, AB2 time
    march (alpha_U = 0: method 4, explicit diffusion; 1: method 6, backward
    Euler), pressure projection; with B, low-Rem induction with the CT source,
    implicit diffusion, phi cleaning, J and J x B.  U, p, B, phi are the
    caller's fields (with their BCs) and are advanced in place.
    ISPs are (iter_max, tol_rel, tol_abs, n_skip_check_res); mpg_dir 1..3
    drives the flow by a mean pressure gradient of mpg_scale; prec_p and
    prec_phi pick the PPE's and the phi cleaning's preconditioner ("Lap_SF",
    the reference's Jacobi, or "MG_FP32")."""

    FIELDS = {"F": 1, "Fnm1": 2, "Ustar": 3, "U_E": 4, "J": 5, "jCrossB": 6, "FB": 7, "ppe_rhs": 8}

    def __init__(self, m: mesh, U, p: SF, B=None, phi: SF | None = None, Re=400.0, Ha=0.0, Rem=1.0, dt=1e-3,
                 alpha_U=1.0, B0=(0.0, 1.0, 0.0), isp_U=(5, 1e-6, 1e-13, 1), isp_p=(5, 1e-6, 1e-13, 1),
                 isp_B=(5, 1e-6, 1e-13, 1), isp_phi=(5, 1e-6, 1e-13, 1), prec_p="Lap_SF", m_ind: mesh | None = None,
                 domain: "mesh_domain | None" = None, sigma_inv=None, mpg_dir: int = 0, mpg_scale: float = 1.0,
                 finite_Rem: bool = False, prec_phi: str = "Lap_SF"):
        P = MhdParams()
        P.Re, P.Ha, P.Rem, P.dt, P.alpha_U = Re, Ha, Rem, dt, alpha_U
        # mean pressure gradient forcing along mpg_dir (1..3, 0 off): compute_add_MPG
        P.mpg_dir, P.mpg_scale = int(mpg_dir), float(mpg_scale)
        P.finite_Rem = 1 if finite_Rem else 0  # Btot = B0 + B (else B0)
        P.induction = 1 if B is not None else 0
        P.prec_p = PCG_solver_SF.PRECS[prec_p]
        # the phi cleaning's preconditioner: the reference's Jacobi, or the mixed multigrid
        P.prec_phi = PCG_solver_SF.PRECS[prec_phi]
        for d in range(3):
            P.B0[d] = B0[d]
        for q, isp in enumerate((isp_U, isp_p, isp_B, isp_phi)):
            P.iter_max[q], P.tol_rel[q], P.tol_abs[q], P.n_skip[q] = isp
        h = C.c_void_p()
        # m_ind, domain: the induction on its own mesh (B, phi, sigma_inv there), the fluid mesh inside it
        check(load().moons_mhd_create_md(m.handle, m_ind.handle if m_ind is not None else None,
                                         domain.handle if domain is not None else None, C.byref(P), _arr(U), p.handle,
                                         _arr(B) if B is not None else None, phi.handle if phi is not None else None,
                                         _arr(sigma_inv) if sigma_inv is not None else None, C.byref(h)))
        self.handle = h
        self.m = m
        self.m_ind = m_ind if m_ind is not None else m
        self._keep = (U, p, B, phi, m_ind, domain, sigma_inv)

    def step(self, n: int = 1):
        info = MhdInfo()
        check(load().moons_mhd_step(self.handle, n, C.byref(info)))
        return {"iterations": dict(zip(("U", "p", "B", "phi"), list(info.iterations))), "ms": info.ms}

    def capture(self) -> None:
        """Capture one step into a HIP graph (solves run exactly their iter_max iterations)."""
        check(load().moons_mhd_graph_capture(self.handle))

    def launch(self, n: int = 1, stream=None) -> None:
        """Replay the captured step n times on `stream` (a Stream, or the library stream)."""
        check(load().moons_mhd_graph_launch(self.handle, n, stream.handle if stream is not None else None))

    def set_buoyancy(self, T: "SF | None", gravity, buoyancy_scale: float | None = None,
                     gravity_scale: float | None = None) -> None:
        """Buoyancy F += cellCenter2Face(T g) * buoyancy_scale and gravity
        F += cellCenter2Face(g) * gravity_scale in the momentum sources, after
        J x B (compute_add_buoyancy / compute_add_gravity,
        code/handwritten/solvers/momentum/momentum_sources.f90:133-167).  g is
        three cell-centred fields; a scale of None switches its term off; T and
        g stay the caller's and are read at every step."""
        self._buoy = (T, list(gravity) if gravity is not None else None)  # keep the fields alive
        gv = (C.c_void_p * 3)(*[f.handle for f in gravity]) if gravity is not None else None
        check(load().moons_mhd_set_buoyancy(self.handle, T.handle if T is not None else None, gv,
                                            0 if buoyancy_scale is None else 1,
                                            0.0 if buoyancy_scale is None else float(buoyancy_scale),
                                            0 if gravity_scale is None else 1,
                                            0.0 if gravity_scale is None else float(gravity_scale)))

    def field(self, which: str):
        out = (C.c_void_p * 6)()
        check(load().moons_mhd_field(self.handle, self.FIELDS[which], out))
        n = 6 if which == "U_E" else (1 if which == "ppe_rhs" else 3)
        mm = self.m_ind if which in ("J", "jCrossB", "FB") else self.m
        return [SF.borrow(mm, out[q]) for q in range(n)]

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                load().moons_mhd_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class Energy_step:
    """The energy equation's step (solve_energy after add_all_energy_sources,
    code/handwritten/governing_equations/energy_extend.f90:236-262): sources
    advection (-div(face(T) U) times advection_scale), diffusion
    (lapUniformCoeff(T) times diffusion_scale; into L when the RK is active),
    KE diffusion, viscous dissipation, Joule heating and volumetric heating
    (code/handwritten/solvers/energy/energy_sources.f90:50-113), each a
    (flag, scale) pair, then method 1 (Euler), 2 (explicit AB2), 3 (AB2 with
    backward-Euler I + c div(k grad), c = dt coeff_implicit), 4 (BDF2 with
    AB2 sources) or 5 (the RK stage's Euler, rk_stages = 3 for the 3-stage
    RK).  T (with BCs) is advanced in place; U is read at each step.  k: three
    face SFs bound to the implicit operator; None gives zeros, as the
    reference's init binds (its operator is then I).  J (x, y, z edge SFs) and
    sigma_inv (cell-centred) feed Joule heating (set_joule)."""

    FIELDS = {"F": 1, "Fnm1": 2, "Tnm1": 3, "L": 4}

    def __init__(self, m: mesh, T: SF, U=None, k=None, dt=1e-3, method=3, Pe=100.0, coeff_implicit=None,
                 advection=(True, -1.0), diffusion=(False, None), isp_T=(5, 1e-6, 1e-13, 1),
                 KE_diffusion=(False, 1.0), viscous_dissipation=(False, 1.0), joule_heating=(False, 1.0),
                 volumetric_heating=(False, 1.0), rk_stages=1, J=None, sigma_inv=None):
        P = NrgParams()
        P.dt, P.method = dt, method
        P.coeff_implicit = -1.0 / Pe if coeff_implicit is None else coeff_implicit
        P.advection, P.advection_scale = (1 if advection[0] else 0), advection[1]
        P.diffusion = 1 if diffusion[0] else 0
        P.diffusion_scale = (1.0 / Pe) if diffusion[1] is None else diffusion[1]
        P.iter_max, P.tol_rel, P.tol_abs, P.n_skip = isp_T
        P.KE_diffusion, P.KE_diffusion_scale = (1 if KE_diffusion[0] else 0), KE_diffusion[1]
        P.viscous_dissipation = 1 if viscous_dissipation[0] else 0
        P.viscous_dissipation_scale = viscous_dissipation[1]
        P.joule_heating, P.joule_heating_scale = (1 if joule_heating[0] else 0), joule_heating[1]
        P.volumetric_heating = 1 if volumetric_heating[0] else 0
        P.volumetric_heating_scale = volumetric_heating[1]
        P.rk_stages = rk_stages
        if k is None:
            k = [SF(m, 1 << d).fill(0.0) for d in range(3)]
        h = C.c_void_p()
        check(load().moons_nrg_create(m.handle, C.byref(P), T.handle, _arr(U) if U is not None else None, _arr(k),
                                      C.byref(h)))
        self.handle = h
        self.m = m
        self._keep = (T, U, k)
        self._joule = None
        if J is not None:
            self.set_joule(J, sigma_inv)

    def set_joule(self, J, sigma_inv: SF):
        """Binds the current density (x, y, z edge SFs) and sigmaInv_CC read by Joule heating."""
        check(load().moons_nrg_set_joule(self.handle, _arr(J), sigma_inv.handle))
        self._joule = (J, sigma_inv)
        return self

    def step(self, n: int = 1):
        info = NrgInfo()
        check(load().moons_nrg_step(self.handle, n, C.byref(info)))
        return {"iterations": info.iterations, "ms": info.ms}

    def field(self, which: str) -> SF:
        out = C.c_void_p()
        check(load().moons_nrg_field(self.handle, self.FIELDS[which], C.byref(out)))
        return SF.borrow(self.m, out.value)

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                load().moons_nrg_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class UNet:
    """Learned PPE preconditioner: 3-level 3D U-net (16/32/64 channels, 3x3x3
    convolutions), BF16 weights/activations with FP32 accumulation on the MFMA
    units, trained on the device by the library's forward/backward kernels.

    Arrays are FP32 with the padded column-major shape of a cell-centred field:
    inputs (4, n0+2, n1+2, n2+2) channel-major, targets/outputs (n0+2, n1+2, n2+2)."""

    def __init__(self, seed: int = 0):
        if not _initialised:
            init(0)
        h = C.c_void_p()
        check(load().moons_unet_create(seed, C.byref(h)))
        self.handle = h
        self.nparam = int(load().moons_unet_num_params(h))

    def get_weights(self) -> np.ndarray:
        w = np.empty(self.nparam, dtype=np.float32)
        check(load().moons_unet_get_weights(self.handle, _fp(w), w.size))
        return w

    def set_weights(self, w: np.ndarray) -> None:
        w = np.ascontiguousarray(w, dtype=np.float32)
        check(load().moons_unet_set_weights(self.handle, _fp(w), w.size))

    @staticmethod
    def _flat_in(in4):
        a = np.asarray(in4, dtype=np.float32)
        return np.ascontiguousarray(np.stack([np.asfortranarray(a[c]).ravel(order="F") for c in range(a.shape[0])]))

    def forward(self, in4: np.ndarray) -> np.ndarray:
        n = [d - 2 for d in in4.shape[1:]]
        x = self._flat_in(in4)
        out = np.zeros(int(np.prod(in4.shape[1:])), dtype=np.float32)
        check(load().moons_unet_forward_host(self.handle, n[0], n[1], n[2], _fp(x), _fp(out)))
        return out.reshape(in4.shape[1:], order="F")

    def grad(self, in4: np.ndarray, target: np.ndarray):
        n = [d - 2 for d in in4.shape[1:]]
        x = self._flat_in(in4)
        t = np.ascontiguousarray(np.asarray(target, dtype=np.float32).ravel(order="F"))
        g = np.empty(self.nparam, dtype=np.float32)
        loss = C.c_double()
        check(load().moons_unet_grad_host(self.handle, n[0], n[1], n[2], _fp(x), _fp(t), C.byref(loss), _fp(g)))
        return loss.value, g

    def train(self, in4_list, target_list, epochs=5, batch=4, lr=1e-3, seed=0) -> np.ndarray:
        n = [d - 2 for d in in4_list[0].shape[1:]]
        X = np.ascontiguousarray(np.concatenate([self._flat_in(a).ravel() for a in in4_list]).astype(np.float32))
        T = np.ascontiguousarray(np.concatenate([np.asarray(t, dtype=np.float32).ravel(order="F")
                                                 for t in target_list]))
        hist = np.zeros(epochs)
        check(load().moons_unet_train(self.handle, n[0], n[1], n[2], len(in4_list), _fp(X), _fp(T), epochs, batch,
                                      float(lr), seed, hist.ctypes.data_as(C.POINTER(C.c_double))))
        return hist

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                load().moons_unet_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def mfma_check(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """D = A @ B (16x16, operands rounded to BF16) through one MFMA with the U-net's lane layout."""
    if not _initialised:
        init(0)
    a = np.ascontiguousarray(A, dtype=np.float32)
    b = np.ascontiguousarray(B, dtype=np.float32)
    d = np.zeros((16, 16), dtype=np.float32)
    check(load().moons_mfma_check(_fp(a), _fp(b), _fp(d)))
    return d


def mfma_check32(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """D = A @ B (A 16x32, B 32x16, operands rounded to BF16) through one
    v_mfma_f32_16x16x32_bf16 with the finest-level convolutions' lane layout."""
    if not _initialised:
        init(0)
    a = np.ascontiguousarray(A, dtype=np.float32)
    b = np.ascontiguousarray(B, dtype=np.float32)
    assert a.shape == (16, 32) and b.shape == (32, 16)
    d = np.zeros((16, 16), dtype=np.float32)
    check(load().moons_mfma_check32(_fp(a), _fp(b), _fp(d)))
    return d


class Stream:
    """A HIP stream of the library (default priority)."""

    def __init__(self):
        if not _initialised:
            init(0)
        h = C.c_void_p()
        check(load().moons_stream_create(C.byref(h)))
        self.handle = h

    def synchronize(self) -> None:
        check(load().moons_stream_synchronize(self.handle))

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                load().moons_stream_destroy(self.handle)
                self.handle = None
        except Exception:
            pass
