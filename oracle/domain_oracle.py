"""CPU restatement of the multi-domain pieces of the induction solve: the
uniform mesh extension, the coordinate overlap of two meshes, embedding and
extraction of fields between them, and the conductivity of the extended
mesh.  TEST INFRASTRUCTURE ONLY: the checker for the device path.

  ext_uniform_IO     code/handwritten/grid/grid_continue.f90:41-47,94-107,211-222
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
