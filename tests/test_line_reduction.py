"""The zebra line kernel's solve order, emulated in float64 on the CPU: C
cells a thread (cyclic reduction inside the thread, one exchange of the
predecessor's eliminated rows), parallel cyclic reduction over the NP / C
threads of a line, then back-substitution widest stride first.  Every
(NP, C) the library instantiates, and line lengths below NP padded with
identity rows, must reproduce a dense solve of the normalised tridiagonal
system."""
import numpy as np
import pytest


def emulate(a, c, d, NP, C):
    """x of rows a x[t-1] + x[t] + c x[t+1] = d (unit diagonal), len(a) <= NP,
    by the kernel's order of operations."""
    nl = len(a)
    A = np.zeros(NP); Cc = np.zeros(NP); D = np.zeros(NP)
    A[:nl], Cc[:nl], D[:nl] = a, c, d
    NQ = NP // C
    A = A.reshape(NQ, C).copy(); Cc = Cc.reshape(NQ, C).copy(); D = D.reshape(NQ, C).copy()
    h = 1
    while h < C:  # local strides, rows t > 0
        for t in range(2 * h, C, 2 * h):
            rr = 1.0 / (1.0 - A[:, t] * Cc[:, t - h] - Cc[:, t] * A[:, t + h])
            an, cn = -A[:, t] * A[:, t - h] * rr, -Cc[:, t] * Cc[:, t + h] * rr
            D[:, t] = (D[:, t] - A[:, t] * D[:, t - h] - Cc[:, t] * D[:, t + h]) * rr
            A[:, t], Cc[:, t] = an, cn
        h *= 2
    a0, c0, d0 = A[:, 0].copy(), Cc[:, 0].copy(), D[:, 0].copy()
    h = 1
    while h < C:  # row 0 with the predecessor's row C - h
        am = np.r_[0.0, A[:-1, C - h]]; cm = np.r_[0.0, Cc[:-1, C - h]]; dm = np.r_[0.0, D[:-1, C - h]]
        rr = 1.0 / (1.0 - a0 * cm - c0 * A[:, h])
        an, cn = -a0 * am * rr, -c0 * Cc[:, h] * rr
        d0 = (d0 - a0 * dm - c0 * D[:, h]) * rr
        a0, c0 = an, cn
        h *= 2
    st = 1
    while st < NQ:  # parallel cyclic reduction over the threads
        z = np.zeros(st)
        am, cm, dm = np.r_[z, a0[:-st]], np.r_[z, c0[:-st]], np.r_[z, d0[:-st]]
        ap, cp, dp = np.r_[a0[st:], z], np.r_[c0[st:], z], np.r_[d0[st:], z]
        rr = 1.0 / (1.0 - a0 * cm - c0 * ap)
        an, cn = -a0 * am * rr, -c0 * cp * rr
        d0 = (d0 - a0 * dm - c0 * dp) * rr
        a0, c0 = an, cn
        st *= 2
    X = np.zeros((NQ, C + 1))
    X[:, 0] = d0
    X[:, C] = np.r_[d0[1:], 0.0]
    h = C // 2
    while h >= 1:
        for t in range(h, C, 2 * h):
            X[:, t] = D[:, t] - A[:, t] * X[:, t - h] - Cc[:, t] * X[:, t + h]
        h //= 2
    return X[:, :C].reshape(-1)[:nl]


@pytest.mark.parametrize("NP,C", [(16, 4), (32, 4), (64, 4), (64, 8), (128, 4), (128, 8), (128, 16)])
def test_line_order_matches_dense_solve(NP, C):
    rng = np.random.default_rng(NP * 31 + C)
    for nl in sorted({4, C + 1, NP // 2 + 3, NP - 1, NP}):
        if nl > NP:
            continue
        # rows like the multigrid's: off-diagonals of one sign, |a| + |c| < 1,
        # strong anisotropy at the ends (stretched walls)
        w = rng.uniform(0.05, 0.49, size=(2, nl))
        w[:, :3] *= 1.9
        a, c = -w[0], -w[1]
        a[0] = 0.0
        c[-1] = 0.0
        d = rng.standard_normal(nl)
        M = np.eye(nl) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
        ref = np.linalg.solve(M, d)
        x = emulate(a, c, d, NP, C)
        assert np.max(np.abs(x - ref)) <= 1e-12 * np.max(np.abs(ref)), (NP, C, nl)
