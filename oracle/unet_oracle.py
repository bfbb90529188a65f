"""NumPy restatement of the learned preconditioner's network: forward pass,
MSE loss and backward pass to every parameter.

TEST INFRASTRUCTURE ONLY: the checker of the device U-net kernels.

Layers (channels): in 4 -> 16 -> 16 (e1) | pool | 32 -> 32 (e2) | pool |
64 -> 64 (b) | up, cat e2 (96) -> 32 -> 32 | up, cat e1 (48) -> 16 -> 16 |
out 16 -> 1.  3x3x3 "same" convolutions with zero padding, ReLU after every
convolution but the last, 2x2x2 average pooling, nearest upsampling.
Weights are stored layer by layer as W[tap][c_in][c_out] then bias[c_out],
tap = (dz+1)*9 + (dy+1)*3 + (dx+1).

With bf16=True the values the device stores in BF16 are rounded the same way
(round to nearest even): weights of every layer, the input, every activation
and every activation gradient; sums are exact (float64) here and FP32 on the
device, so the two agree to BF16 rounding flips.  With bf16=False everything
is float64 (used to pin the backward pass against finite differences).
"""
from __future__ import annotations

import numpy as np

LAYERS = [(4, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (96, 32), (32, 32), (48, 16), (16, 16), (16, 1)]


def n_params():
    return sum(27 * ci * co + co for ci, co in LAYERS)


def unpack(w):
    out, off = [], 0
    for ci, co in LAYERS:
        W = np.asarray(w[off:off + 27 * ci * co], dtype=np.float64).reshape(27, ci, co)
        off += 27 * ci * co
        b = np.asarray(w[off:off + co], dtype=np.float64)
        off += co
        out.append((W, b))
    return out


def pack(grads):
    return np.concatenate([np.concatenate([dW.ravel(), db.ravel()]) for dW, db in grads])


def bf16(x):
    """Round to BF16 (nearest even) and back, via float32."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def pad(a):
    """(n2, n1, n0, C) -> zero halo of one cell."""
    return np.pad(a, ((1, 1), (1, 1), (1, 1), (0, 0)))


def conv(Xp, W, b):
    """Xp: padded (n2+2, n1+2, n0+2, Cin); returns unpadded (n2, n1, n0, Cout) pre-activation."""
    n2, n1, n0 = (s - 2 for s in Xp.shape[:3])
    Y = np.zeros((n2, n1, n0, W.shape[2]))
    for t, (dz, dy, dx) in enumerate(TAPS):
        Y += Xp[1 + dz:1 + dz + n2, 1 + dy:1 + dy + n1, 1 + dx:1 + dx + n0, :] @ W[t]
    return Y + b


def conv_T(dYp, W):
    """gradient w.r.t. the input of conv: dX[u] = sum_tap W[tap] dY[u - tap]."""
    n2, n1, n0 = (s - 2 for s in dYp.shape[:3])
    dX = np.zeros((n2, n1, n0, W.shape[1]))
    for t, (dz, dy, dx) in enumerate(TAPS):
        dX += dYp[1 - dz:1 - dz + n2, 1 - dy:1 - dy + n1, 1 - dx:1 - dx + n0, :] @ W[t].T
    return dX


def conv_dW(Xp, dY):
    n2, n1, n0 = dY.shape[:3]
    dW = np.zeros((27, Xp.shape[3], dY.shape[3]))
    flat = dY.reshape(-1, dY.shape[3])
    for t, (dz, dy, dx) in enumerate(TAPS):
        xs = Xp[1 + dz:1 + dz + n2, 1 + dy:1 + dy + n1, 1 + dx:1 + dx + n0, :].reshape(-1, Xp.shape[3])
        dW[t] = xs.T @ flat
    return dW, flat.sum(axis=0)


def pool(a):
    n2, n1, n0, C = a.shape
    return a.reshape(n2 // 2, 2, n1 // 2, 2, n0 // 2, 2, C).sum(axis=(1, 3, 5)) * 0.125


def up(a):
    return a.repeat(2, axis=0).repeat(2, axis=1).repeat(2, axis=2)


class Net:
    def __init__(self, w, bf=True):
        self.bf = bf
        self.P = unpack(w)
        r = bf16 if bf else (lambda x: np.asarray(x, dtype=np.float64))
        self.r = r
        self.Wq = [(r(W), b) for W, b in self.P]  # BF16 weights, FP32 biases

    def forward(self, x):
        """x: (n2, n1, n0, 4) unpadded input; returns (n2, n1, n0) output and keeps the activations."""
        r, Wq = self.r, self.Wq
        relu = lambda a: np.maximum(a, 0.0)
        A = {}
        A["in"] = pad(r(x))
        A["a0"] = pad(r(relu(conv(A["in"], *Wq[0]))))
        A["e1"] = pad(r(relu(conv(A["a0"], *Wq[1]))))
        A["p1"] = pad(r(pool(A["e1"][1:-1, 1:-1, 1:-1])))
        A["a2"] = pad(r(relu(conv(A["p1"], *Wq[2]))))
        A["e2"] = pad(r(relu(conv(A["a2"], *Wq[3]))))
        A["p2"] = pad(r(pool(A["e2"][1:-1, 1:-1, 1:-1])))
        A["a4"] = pad(r(relu(conv(A["p2"], *Wq[4]))))
        A["b"] = pad(r(relu(conv(A["a4"], *Wq[5]))))
        A["c2"] = pad(np.concatenate([up(A["b"][1:-1, 1:-1, 1:-1]), A["e2"][1:-1, 1:-1, 1:-1]], axis=3))
        A["a6"] = pad(r(relu(conv(A["c2"], *Wq[6]))))
        A["d2"] = pad(r(relu(conv(A["a6"], *Wq[7]))))
        A["c1"] = pad(np.concatenate([up(A["d2"][1:-1, 1:-1, 1:-1]), A["e1"][1:-1, 1:-1, 1:-1]], axis=3))
        A["a8"] = pad(r(relu(conv(A["c1"], *Wq[8]))))
        A["d1"] = pad(r(relu(conv(A["a8"], *Wq[9]))))
        out = conv(A["d1"], *Wq[10])[..., 0]
        self.A = A
        return out

    def backward(self, dout):
        """dout: (n2, n1, n0) gradient of the loss w.r.t. the output; returns per-layer (dW, db)."""
        r, Wq, A = self.r, self.Wq, self.A
        G = [None] * 11
        inner = lambda a: a[1:-1, 1:-1, 1:-1]
        mask = lambda g, act: g * (inner(act) > 0)
        G[10] = conv_dW(A["d1"], dout[..., None])
        g = r(conv_T(pad(dout[..., None]), Wq[10][0]))  # d d1
        g = mask(g, A["d1"]); G[9] = conv_dW(A["a8"], g); g = r(conv_T(pad(g), Wq[9][0]))
        g = mask(g, A["a8"]); G[8] = conv_dW(A["c1"], g); g = r(conv_T(pad(g), Wq[8][0]))  # d c1 (48)
        n2, n1, n0 = g.shape[:3]
        g_d2 = r(g[..., :32].reshape(n2 // 2, 2, n1 // 2, 2, n0 // 2, 2, 32).sum(axis=(1, 3, 5)))
        g_e1 = g[..., 32:]
        g = mask(g_d2, A["d2"]); G[7] = conv_dW(A["a6"], g); g = r(conv_T(pad(g), Wq[7][0]))
        g = mask(g, A["a6"]); G[6] = conv_dW(A["c2"], g); g = r(conv_T(pad(g), Wq[6][0]))  # d c2 (96)
        h2, h1, h0 = g.shape[:3]
        g_b = r(g[..., :64].reshape(h2 // 2, 2, h1 // 2, 2, h0 // 2, 2, 64).sum(axis=(1, 3, 5)))
        g_e2 = g[..., 64:]
        g = mask(g_b, A["b"]); G[5] = conv_dW(A["a4"], g); g = r(conv_T(pad(g), Wq[5][0]))
        g = mask(g, A["a4"]); G[4] = conv_dW(A["p2"], g); g = r(conv_T(pad(g), Wq[4][0]))  # d p2
        g_e2 = r(g_e2 + 0.125 * up(g))
        g = mask(g_e2, A["e2"]); G[3] = conv_dW(A["a2"], g); g = r(conv_T(pad(g), Wq[3][0]))
        g = mask(g, A["a2"]); G[2] = conv_dW(A["p1"], g); g = r(conv_T(pad(g), Wq[2][0]))  # d p1
        g_e1 = r(g_e1 + 0.125 * up(g))
        g = mask(g_e1, A["e1"]); G[1] = conv_dW(A["a0"], g); g = r(conv_T(pad(g), Wq[1][0]))
        g = mask(g, A["a0"]); G[0] = conv_dW(A["in"], g)
        return G

    def loss_grad(self, x, target):
        out = self.forward(x)
        nvox = out.size
        d = out - target
        loss = float((d * d).sum() / nvox)
        G = self.backward(2.0 * d / nvox)
        return loss, pack(G), out


def to_internal(in4_padded_F):
    """(4, n0+2, n1+2, n2+2) channel-major padded (column-major per channel) -> (n2, n1, n0, 4) interior."""
    a = np.asarray(in4_padded_F, dtype=np.float64)[:, 1:-1, 1:-1, 1:-1]
    return np.transpose(a, (3, 2, 1, 0))


def from_internal(out):
    """(n2, n1, n0) -> (n0, n1, n2) (interior of a padded column-major field)."""
    return np.transpose(out, (2, 1, 0))
