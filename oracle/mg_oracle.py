"""FP64 NumPy restatement of the multigrid V-cycle of moons_amd/csrc/mg.hip.

TEST INFRASTRUCTURE ONLY.  The reference compiles no multigrid (its V-cycle
is dead code: code_not_compiled/9.6.2017/MG/MG.f90:241-396 -- smooth,
residual, restrict, recurse, prolongate, correct, smooth), so this file
restates the build's own V-cycle step for step, from the same mathematics:
node-coarsened levels (code/handwritten/grid/coordinates_extend.f90:336-355),
the Laplacian with homogeneous BCs folded into per-direction weights,
red-black Gauss-Seidel -- or, on grid levels stretched along a non-periodic
y or z direction, zebra line Gauss-Seidel along those directions --,
volume-weighted restriction of residuals, trilinear prolongation in true
coordinates.  Pinned by: the level operator equals the
oracle's vol-weighted Laplacian divided by vol (assembled densely), the V-cycle
is a convergent stationary iteration, and MG-preconditioned CG reaches the
dense solution.
"""
from __future__ import annotations

import numpy as np

PERIODIC, DIRICHLET, NEUMANN, ANTISYM = 3, 1, 2, 5


def _dirichlet_like(t):
    return t in (DIRICHLET, ANTISYM)


def coarsen_nodes(hn):
    phys = hn[1:-1][::2]
    return np.concatenate([[phys[0] - (phys[1] - phys[0])], phys, [phys[-1] + (phys[-1] - phys[-2])]])


class Level:
    def __init__(self, hn3, bct):
        self.hn = hn3
        self.n = [len(h) - 3 for h in hn3]
        self.per = [bct[2 * d] == PERIODIC for d in range(3)]
        self.wm, self.wp, self.dg, self.hc, self.dhn = [], [], [], [], []
        for d in range(3):
            hn = hn3[d]
            n = self.n[d]
            s = n + 2
            dhn = hn[1:] - hn[:-1]
            hc = 0.5 * (hn[1:] + hn[:-1])
            dhc = hc[1:] - hc[:-1]
            wm = np.zeros(s)
            wp = np.zeros(s)
            dg = np.zeros(s)
            bmin, bmax = bct[2 * d], bct[2 * d + 1]
            for i in range(1, n + 1):
                a = 1.0 / (dhc[i - 1] * dhn[i])
                c = 1.0 / (dhc[i] * dhn[i])
                diag = -(a + c)
                if i == 1 and bmin != PERIODIC:
                    diag += -a if _dirichlet_like(bmin) else a
                    a = 0.0
                if i == n and bmax != PERIODIC:
                    diag += -c if _dirichlet_like(bmax) else c
                    c = 0.0
                wm[i], wp[i], dg[i] = a, c, diag
            self.wm.append(wm)
            self.wp.append(wp)
            self.dg.append(dg)
            self.hc.append(hc)
            self.dhn.append(dhn)
        # zebra lines along y / z where the cell widths vary by more than 1.5x,
        # the direction is not periodic, and the level has more than 32^3 cells
        cells = self.n[0] * self.n[1] * self.n[2]
        self.line_dirs = []
        for d in (1, 2):
            w = self.dhn[d][1:self.n[d] + 1]
            if (bct[2 * d] != PERIODIC and 4 <= self.n[d] <= LINE_MAX and cells > LINE_MIN_CELLS
                    and w.max() > 1.5 * w.min()):
                self.line_dirs.append(d)

    def shape(self):
        return tuple(n + 2 for n in self.n)

    def _nb(self, x, d, sign):
        """x shifted to the -/+ neighbour along d on interior cells, with wrap for periodic."""
        n = self.n[d]
        xi = np.moveaxis(x, d, 0)
        idx = np.arange(1, n + 1) + sign
        if self.per[d]:
            idx = np.where(idx == 0, n, idx)
            idx = np.where(idx == n + 1, 1, idx)
        else:
            idx = np.clip(idx, 1, n)
        other = [slice(1, m + 1) for k, m in enumerate(self.n) if k != d]
        v = xi[(idx,) + tuple(other)]
        return np.moveaxis(v, 0, d)

    def _w(self, arr, d):
        n = self.n[d]
        shp = [1, 1, 1]
        shp[d] = n
        return arr[1:n + 1].reshape(shp)

    def offdiag(self, x):
        s = 0.0
        for d in range(3):
            s = s + self._w(self.wm[d], d) * self._nb(x, d, -1) + self._w(self.wp[d], d) * self._nb(x, d, +1)
        return s

    def diag(self):
        return self._w(self.dg[0], 0) + self._w(self.dg[1], 1) + self._w(self.dg[2], 2)

    def interior(self, a):
        return a[1:-1, 1:-1, 1:-1]

    def color_mask(self, c):
        i, j, k = np.meshgrid(*[np.arange(1, n + 1) for n in self.n], indexing="ij")
        return ((i + j + k) % 2) == c

    def smooth(self, x, f, color, zero_guess):
        m = self.color_mask(color)
        fi = self.interior(f)
        if zero_guess:
            new = fi / self.diag()
        else:
            new = (fi - self.offdiag(x)) / self.diag()
        xi = self.interior(x)
        xi[m] = new[m]

    def line_smooth(self, x, f, d, color, zero_guess):
        """Zebra line Gauss-Seidel along d (1 or 2) on the lines of colour c:
        (i + k) % 2 == c for y-lines, (i + j) % 2 == c for z-lines; each line's
        tridiagonal system solved exactly (Thomas), the off-line neighbours on
        the right-hand side (zero with zero_guess)."""
        n = self.n
        fi = self.interior(f)
        s = 0.0
        if not zero_guess:
            for e in range(3):
                if e != d:
                    s = s + self._w(self.wm[e], e) * self._nb(x, e, -1) + self._w(self.wp[e], e) * self._nb(x, e, +1)
        rhs = np.moveaxis(fi - s, d, 0)
        dg = np.moveaxis(self.diag() * np.ones(tuple(n)), d, 0)
        a, c = self.wm[d][1:n[d] + 1], self.wp[d][1:n[d] + 1]
        m = n[d]
        cp = np.zeros_like(rhs)
        gp = np.zeros_like(rhs)
        cprev = np.zeros(rhs.shape[1:])
        gprev = np.zeros(rhs.shape[1:])
        for t in range(m):
            beta = dg[t] - a[t] * cprev
            gp[t] = (rhs[t] - a[t] * gprev) / beta
            cp[t] = c[t] / beta
            cprev, gprev = cp[t], gp[t]
        sol = np.zeros_like(rhs)
        xn = np.zeros(rhs.shape[1:])
        for t in range(m - 1, -1, -1):
            xn = gp[t] - cp[t] * xn
            sol[t] = xn
        sol = np.moveaxis(sol, 0, d)
        others = [e for e in range(3) if e != d]
        idx = np.meshgrid(*[np.arange(1, k + 1) for k in n], indexing="ij")
        mask = ((idx[others[0]] + idx[others[1]]) % 2) == color
        xi = self.interior(x)
        xi[mask] = sol[mask]

    def sweep(self, x, f, forward, zero_first):
        """One smoothing sweep: red-black points (forward: colour 0 first), or
        where the level has line directions, y-lines colour 0, 1 then z-lines
        colour 0, 1 (forward) or the reverse order (backward)."""
        if not self.line_dirs:
            c0 = 0 if forward else 1
            self.smooth(x, f, c0, zero_first)
            self.smooth(x, f, 1 - c0, False)
            return
        if forward:
            for a, d in enumerate(self.line_dirs):
                for c in (0, 1):
                    self.line_smooth(x, f, d, c, zero_first and a == 0 and c == 0)
        else:
            for d in self.line_dirs[::-1]:
                for c in (1, 0):
                    self.line_smooth(x, f, d, c, False)

    def residual(self, x, f):
        return self.interior(f) - (self.offdiag(x) + self.diag() * self.interior(x))


COARSE_CELLS = 4  # a level of at most this many cells per direction is the coarsest
LINE_MIN_CELLS = 32 * 32 * 32  # levels of at most this many cells are smoothed by points
LINE_MAX = 128  # longest line the device holds in LDS


def build(hn3, bct, max_levels=16, coarse_cells=COARSE_CELLS):
    levels = [Level([np.asarray(h, float) for h in hn3], bct)]
    while len(levels) < max_levels:
        L = levels[-1]
        if not all(n % 2 == 0 and n // 2 >= 2 for n in L.n):
            break
        if any(bct[2 * d] == 3 and (L.n[d] // 2) % 2 != 0 for d in range(3)):
            break
        if max(L.n) <= coarse_cells:
            break
        levels.append(Level([coarsen_nodes(h) for h in L.hn], bct))
    return levels


def restrict(F: Level, C: Level, res):
    """coarse f = sum of children (child vol / coarse vol) * residual."""
    rf = []
    for d in range(3):
        dc = C.dhn[d][1:C.n[d] + 1]
        df = F.dhn[d][1:F.n[d] + 1]
        rf.append(df / np.repeat(dc, 2))
    w = rf[0][:, None, None] * rf[1][None, :, None] * rf[2][None, None, :]
    v = w * res
    n0, n1, n2 = C.n
    return v.reshape(n0, 2, n1, 2, n2, 2).sum(axis=(1, 3, 5))


def prolong(F: Level, C: Level, xc, bct):
    """trilinear interpolation of the coarse correction to fine interior cells."""
    idx, wa, wb, idx2 = [], [], [], []
    for d in range(3):
        n, nc = F.n[d], C.n[d]
        I = (np.arange(1, n + 1) + 1) // 2
        left = (np.arange(1, n + 1) % 2) == 1
        I2 = np.where(left, I - 1, I + 1)
        hcf = F.hc[d][1:n + 1]
        hcc = C.hc[d]
        a = (hcf - hcc[I2]) / (hcc[I] - hcc[I2])
        b = 1.0 - a
        bmin, bmax = bct[2 * d], bct[2 * d + 1]
        for q in range(n):
            if I2[q] in (0, nc + 1):
                bt = bmin if left[q] else bmax
                if bt == PERIODIC:
                    I2[q] = nc if I2[q] == 0 else 1
                else:
                    I2[q] = I[q]
                    if _dirichlet_like(bt):
                        b[q] = -b[q]
        idx.append(I)
        idx2.append(I2)
        wa.append(a)
        wb.append(b)
    e = 0.0
    for cx, wx in ((idx[0], wa[0]), (idx2[0], wb[0])):
        for cy, wy in ((idx[1], wa[1]), (idx2[1], wb[1])):
            for cz, wz in ((idx[2], wa[2]), (idx2[2], wb[2])):
                e = e + wx[:, None, None] * wy[None, :, None] * wz[None, None, :] * xc[np.ix_(cx, cy, cz)]
    return e


def vcycle(levels, f0, bct, nu1=2, nu2=2, coarse_sweeps=16, x0=None, gamma_level=-1, gamma_cycles=1):
    """x0 = V f0 (padded arrays; f0 = r/vol on the fine level).  With x0 given,
    the fine level's pre-smoothing starts from it (the learned initial guess)
    instead of from zero.  Level gamma_level's visit runs its smoothing,
    restriction, coarser cycle, prolongation and smoothing gamma_cycles times,
    each from the previous result (a gamma-cycle at that level)."""
    xs = [np.zeros(L.shape()) for L in levels]
    if x0 is not None:
        xs[0] = np.array(x0, dtype=float)
    fs = [None] * len(levels)
    fs[0] = f0

    def cyc(l):
        L = levels[l]
        x, f = xs[l], fs[l]
        if l == len(levels) - 1:
            for s in range(coarse_sweeps):
                L.smooth(x, f, 0, s == 0)
                L.smooth(x, f, 1, False)
            return
        for k in range(gamma_cycles if l == gamma_level else 1):
            for s in range(nu1):
                L.sweep(x, f, True, s == 0 and k == 0 and not (l == 0 and x0 is not None))
            C = levels[l + 1]
            fc = np.zeros(C.shape())
            fc[1:-1, 1:-1, 1:-1] = restrict(L, C, L.residual(x, f))
            fs[l + 1] = fc
            cyc(l + 1)
            L.interior(x)[...] += prolong(L, C, xs[l + 1], bct)
            for s in range(nu2):
                L.sweep(x, f, False, False)

    cyc(0)
    return xs[0]
