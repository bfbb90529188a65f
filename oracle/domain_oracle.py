"""CPU restatement of the multi-domain pieces of the induction solve: the
uniform mesh extension, the coordinate overlap of two meshes, embedding and
extraction of fields between them, and the conductivity of the extended
mesh.  TEST INFRASTRUCTURE ONLY: the checker for the device path.

  ext_uniform_IO     code/handwritten/grid/grid_continue.f90:41-47,94-107,211-222
  ext_Roberts_*, ext_prep/app_Roberts_*
                     code/handwritten/grid/grid_continue.f90:49-77,107-193,224-306 (ext)
                     (uniformDirection_func, code/handwritten/grid/coordinate_distribution_funcs.f90:83-111;
                     prepend_ghost_c / append_ghost_c, code/handwritten/grid/coordinates_extend.f90:540-560)
  overlap            get_C_overlap_OL / get_N_overlap_OL (p = 1, tol 1e-12),
                     code/handwritten/sub_domain/overlap_extend.f90:204-278,364-391;
                     the physical sub-domain, code/handwritten/physical_domain/physical_sub_domain_extend.f90:36-45
  embed / extract    embed_GF / extract_GF over EE_shape's ranges,
                     code/handwritten/grid_field/GF_embed_extract/GF_embed_extract.f90:14-60,
                     code/handwritten/ops/ops_embedExtract.f90:280-320
  sigma_inv          set_sigma_inv_VF, code/handwritten/solvers/induction/induction_aux.f90:178-194
Index ranges here are 0-based and inclusive.
"""
from __future__ import annotations

import numpy as np


def ext_uniform(hn, N: int) -> np.ndarray:
    """Nodes (with ghosts) extended by N uniform cells at both ends: the
    spacing of each end's ghost cell, continued from hmin down and hmax up."""
    hn = [float(v) for v in hn]
    # prepend: hmin = hn(2), dh = dhn(1); uniformLeft(hmin, dh, N) replaces the ghost and hmin
    hmin, dh = hn[1], hn[1] - hn[0]
    left = [hmin + (-1.0 * float(i)) * dh for i in range(N + 1)][::-1]
    a = left + hn[2:]
    a = [a[0] - (a[1] - a[0])] + a
    # append: hmax = hn(sn-1), dh = dhn_e = hn(sn) - hn(sn-1)
    hmax, dh = a[-2], a[-1] - a[-2]
    right = [hmax + (1.0 * float(i)) * dh for i in range(N + 1)]
    a = a[:-2] + right
    a = a + [a[-1] + (a[-1] - a[-2])]
    return np.array(a)


# ---------------------------------------------------------------------------
# Roberts mesh extensions (grid_continue.f90:49-77,107-193,224-306) with the
# stretch-parameter matching of coordinate_stretch_param_match.f90:47-95,176-233
# and the Roberts transformations of coordinate_distribution_funcs.f90:145-248.
# Scalar libm arithmetic (math.pow), in the reference's evaluation order.
# ---------------------------------------------------------------------------
import math

_EPS = 2.0 ** -52  # epsilon(1.0_cp)


def _t2_h(b, a, i, N):
    """T2_h_i (coordinate_stretch_param_match.f90:47-55)."""
    g = (b + 1.0) / (b - 1.0)
    t = (float(i) / float(N) - a) / (1.0 - a)
    gt = math.pow(g, t)
    return ((b + 2.0 * a) * gt - b + 2.0 * a) / ((2.0 * a + 1.0) * (1.0 + gt))


def _t2_prime(b, a, i, N):
    """T2_prime_h_i (:57-65); x**2.0 is x*x as the compiler emits it."""
    g = (b + 1.0) / (b - 1.0)
    t = (float(i) / float(N) - a) / (1.0 - a)
    gt, g2t = math.pow(g, t), math.pow(g, 2.0 * t)
    num = -(b * b) * g2t + b * b + 4.0 * b * gt * t + g2t - 1.0
    return -num / ((2.0 * a + 1.0) * (b - 1.0) * (b + 1.0) * ((gt + 1.0) * (gt + 1.0)))


def _root_near_hmax(b, hmin, hmax, a, N, dh):
    return _t2_h(b, a, N, N) - _t2_h(b, a, N - 1, N) - dh / (hmax - hmin)


def _prime_near_hmax(b, a, N):
    return _t2_prime(b, a, N, N) - _t2_prime(b, a, N - 1, N)


def _root_near_hmin(b, hmin, hmax, a, N, dh):
    return _t2_h(b, a, 1, N) - _t2_h(b, a, 0, N) - dh / (hmax - hmin)


def _prime_near_hmin(b, a, N):
    return _t2_prime(b, a, 1, N) - _t2_prime(b, a, 0, N)


def _newton_t2(root, prime, beta, hmin, hmax, a, N, dh):
    """newtonT2 (:176-203); a non-converged match is an error (the reference
    writes mesh_generation_error and stops at init)."""
    tol = 10.0 * _EPS
    beta_low = 1.0 + math.pow(10.0, -10.0)
    fb = 0.0
    for _ in range(1000000):
        fb = root(beta, hmin, hmax, a, N, dh)
        fp = prime(beta, a, N)
        if abs(fb) < tol:
            break
        if beta > 1000.0:
            break
        if beta < beta_low:
            beta = 1.1
            break
        beta = beta - fb / fp
    if math.isnan(beta) or abs(fb) > tol:
        raise ValueError(f"stretch parameter matching did not converge (beta {beta}, f {fb})")
    return beta


def beta_dh_small(hmin, hmax, N, dh):
    return _newton_t2(_root_near_hmax, _prime_near_hmax, 1.0 + math.pow(10.0, -10.0), hmin, hmax, 0.0, N, dh)


def beta_dh_big(hmin, hmax, N, dh):
    return _newton_t2(_root_near_hmin, _prime_near_hmin, 1.0 + math.pow(10.0, -10.0), hmin, hmax, 0.0, N, dh)


def beta_dh_both(hmin, hmax, N, dh):
    return _newton_t2(_root_near_hmax, _prime_near_hmax, 1.0 + math.pow(10.0, -10.0), hmin, hmax, 0.5, N, dh)


def _normalised(hmin, hmax, N):
    dh = (hmax - hmin) / float(N)
    return [(hmin + float(i) * dh - hmin) / (hmax - hmin) for i in range(N + 1)]


def _check_ratios(h, increasing):
    for i in range(len(h) - 2):
        R = (h[i + 2] - h[i + 1]) / (h[i + 1] - h[i])
        if (increasing and R < 1.0) or (not increasing and R > 1.0):
            raise ValueError("Roberts coordinates fail the spacing-ratio check")


def roberts_left(hmin, hmax, N, beta):
    """robertsLeft_func = transformation1 (:145-190)."""
    a, b = 1.0, beta
    g = (b + 1.0) / (b - 1.0)
    hn = []
    for x in _normalised(hmin, hmax, N):
        e = math.pow(g, a - x)
        hn.append(((b + a) - (b - a) * e) / (e + a))
    _check_ratios(hn, True)
    return [hmin + (hmax - hmin) * v for v in hn]


def _transformation2(hmin, hmax, N, a, beta):
    b = beta
    g = (b + 1.0) / (b - 1.0)
    hn = []
    for x in _normalised(hmin, hmax, N):
        e = math.pow(g, (x - a) / (1.0 - a))
        hn.append(((b + 2.0 * a) * e - b + 2.0 * a) / ((2.0 * a + 1.0) * (1.0 + e)))
    if abs(a) < math.pow(10.0, -10.0):
        _check_ratios(hn, False)
    return [hmin + (hmax - hmin) * v for v in hn]


def roberts_right(hmin, hmax, N, beta):
    """robertsRight_func = transformation2, alpha = 0 (:192-248, 317-323)."""
    return _transformation2(hmin, hmax, N, 0.0, beta)


def roberts_both(hmin, hmax, N, beta):
    """robertsBoth_func = transformation2, alpha = 1/2 (:325-331)."""
    return _transformation2(hmin, hmax, N, 0.5, beta)


def _needs_more_points(b, max_ratio):
    """needs_more_points_func (coordinate_distribution_funcs_iterate.f90:136-151)."""
    if len(b) <= 2:
        return False
    d = [b[i + 1] - b[i] for i in range(len(b) - 1)]
    R = [max(d[i] / d[i + 1], d[i + 1] / d[i]) for i in range(len(d) - 1)]
    return max(R) > max_ratio


_MATCH = {"small": beta_dh_small, "big": beta_dh_big, "both": beta_dh_both}
_DIST = {"L": roberts_left, "R": roberts_right, "B": roberts_both}


def _segment(lo, hi, N, dh, match, dist, n_iter, max_ratio):
    """The N_iter loop of ext_prep/app_Roberts_*: N_final = b%N (the points
    of the last segment) and one more when the segment fails the quality check."""
    N_final = N
    b = None
    for _ in range(n_iter):
        beta = _MATCH[match](lo, hi, N_final, dh)
        b = _DIST[dist](lo, hi, N_final, beta)
        N_final = len(b)
        if _needs_more_points(b, max_ratio):
            N_final = N_final + 1
        else:
            break
    return b


def _prepend(hn, b):
    """process_prepend: drop the ghost and hmin, put b in front, a new ghost."""
    a = list(b) + list(hn[2:])
    return [a[0] - (a[1] - a[0])] + a


def _append(hn, b):
    """process_append: drop hmax and the ghost, put b behind, a new ghost."""
    a = list(hn[:-2]) + list(b)
    return a + [a[-1] + (a[-1] - a[-2])]


EXT_KINDS = ("uniform", "near", "far", "B", "prep_uniform", "prep_L", "prep_R", "prep_B", "prep_C2F",
             "app_uniform", "app_L", "app_R", "app_B", "app_C2F")


def ext_prep(hn, kind, L, N, n_iter=1, max_ratio=1.0):
    hn = [float(v) for v in hn]
    hmin, dh = hn[1], hn[1] - hn[0]
    if kind == "uniform":
        b = [hmin + (-1.0 * float(i)) * dh for i in range(N + 1)][::-1]
    else:
        match, dist = {"L": ("small", "L"), "R": ("small", "R"), "B": ("both", "B"), "C2F": ("big", "L")}[kind]
        b = _segment(hmin - L, hmin, N, dh, match, dist, n_iter, max_ratio)
    return _prepend(hn, b)


def ext_app(hn, kind, L, N, n_iter=1, max_ratio=1.0):
    hn = [float(v) for v in hn]
    hmax, dh = hn[-2], hn[-1] - hn[-2]
    if kind == "uniform":
        b = [hmax + (1.0 * float(i)) * dh for i in range(N + 1)]
    else:
        match, dist = {"L": ("small", "L"), "R": ("small", "R"), "B": ("both", "B"), "C2F": ("big", "R")}[kind]
        b = _segment(hmax, hmax + L, N, dh, match, dist, n_iter, max_ratio)
    return _append(hn, b)


def ext(hn, kind, L=0.0, N=1, n_iter=1, max_ratio=1.0) -> np.ndarray:
    """The grid_continue extension `kind` (EXT_KINDS) of the node array hn
    (ghosts included) by segments of N cells and length L (uniform kinds
    ignore L).  n_iter, max_ratio: the mesh-quality parameters (N_iter and
    max_mesh_stretch_ratio; 1 iteration for a manual mesh)."""
    if kind == "uniform":
        out = ext_app(ext_prep(hn, "uniform", L, N), "uniform", L, N)
    elif kind == "near":
        out = ext_app(ext_prep(hn, "R", L, N, n_iter, max_ratio), "L", L, N, n_iter, max_ratio)
    elif kind == "far":
        out = ext_app(ext_prep(hn, "L", L, N, n_iter, max_ratio), "R", L, N, n_iter, max_ratio)
    elif kind == "B":
        out = ext_app(ext_prep(hn, "B", L, N, n_iter, max_ratio), "B", L, N, n_iter, max_ratio)
    elif kind.startswith("prep_"):
        out = ext_prep(hn, kind[5:], L, N, n_iter, max_ratio)
    elif kind.startswith("app_"):
        out = ext_app(hn, kind[4:], L, N, n_iter, max_ratio)
    else:
        raise ValueError(kind)
    return np.array(out)


def _inside(x, h, tol):
    return abs(x - h) < tol


def overlap(h1, h2, tol=1e-12, pad=1):
    """(i1 lo, i1 hi, i2 lo, i2 hi): points of h1 and h2 (cell centres or
    nodes of the two meshes along one direction) that coincide within tol,
    the pad points at each end of both excluded; None when none do."""
    n1, n2 = len(h1), len(h2)
    p1 = 0 if n1 - pad <= 0 or pad <= 0 else pad
    p2 = 0 if n2 - pad <= 0 or pad <= 0 else pad
    hits = [(i, j) for i in range(p1, n1 - p1) for j in range(p2, n2 - p2) if _inside(h2[j], h1[i], tol)]
    if not hits:
        return None
    i = [q[0] for q in hits]
    j = [q[1] for q in hits]
    return (min(i), max(i), min(j), max(j))


class Domain:
    """Where the interior mesh (hn_int per direction) sits in the total mesh
    (hn_tot): per direction the cell-centre (C) and node (N) overlaps."""

    def __init__(self, hn_int, hn_tot):
        self.C, self.N = [], []
        for a, b in zip(hn_int, hn_tot):
            a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
            hc_a, hc_b = (a[1:] + a[:-1]) / 2.0, (b[1:] + b[:-1]) / 2.0
            self.C.append(overlap(hc_a, hc_b))
            self.N.append(overlap(a, b))

    def ranges(self, dl):
        """Slices into the interior and the total arrays of location dl."""
        si, st = [], []
        for d in range(3):
            r = self.N[d] if (dl >> d) & 1 else self.C[d]
            si.append(slice(r[0], r[1] + 1))
            st.append(slice(r[2], r[3] + 1))
        return tuple(si), tuple(st)


def embed(total, interior, dl, dom: Domain):
    out = np.array(total, order="F", copy=True)
    si, st = dom.ranges(dl)
    out[st] = interior[si]
    return out


def extract(interior, total, dl, dom: Domain):
    out = np.array(interior, order="F", copy=True)
    si, st = dom.ranges(dl)
    out[si] = total[st]
    return out


def sigma_inv(shape, dl, dom: Domain, sig_local_over_sig_f):
    """1/ratio everywhere, 1 where the conducting sub-mesh lies."""
    s = np.full(shape, 1.0 / sig_local_over_sig_f, order="F")
    _, st = dom.ranges(dl)
    s[st] = 1.0
    return s
