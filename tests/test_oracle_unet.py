"""Pins the NumPy U-net restatement: its backward pass equals central finite
differences of its forward pass (float64), the BF16 rounding is round-to-
nearest-even, and the parameter count matches the layer table."""
import numpy as np

from oracle import unet_oracle as UO


def test_bf16_rounding_nearest_even():
    x = np.array([1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-9, 1.0 + 2**-9, -2.5, 3.0e-3], dtype=np.float64)
    r = UO.bf16(x)
    assert r[0] == 1.0 and r[4] == -2.5
    # exact BF16 values survive; ties go to even
    assert UO.bf16(np.array([1.0 + 2**-7]))[0] == 1.0 + 2**-7
    assert UO.bf16(np.array([1.0 + 2**-8]))[0] == 1.0          # tie -> even (mantissa 0)
    assert UO.bf16(np.array([1.0 + 3 * 2**-8]))[0] == 1.0 + 2**-6  # tie -> even (mantissa 2)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2**-8)


def test_param_count():
    assert UO.n_params() == sum(27 * a * b + b for a, b in UO.LAYERS)
    assert UO.n_params() == 354_993


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    w = rng.standard_normal(UO.n_params()) * 0.2
    x = rng.standard_normal((4, 4, 4, 4))
    t = rng.standard_normal((4, 4, 4))
    net = UO.Net(w, bf=False)
    loss, g, _ = net.loss_grad(x, t)
    idx = rng.choice(w.size, 40, replace=False)
    # include every bias of the output layer and some first-layer weights
    idx = np.concatenate([idx, [w.size - 1, 0, 5, 27 * 4 * 16 + 3]])
    eps = 1e-6
    for q in idx:
        wp, wm = w.copy(), w.copy()
        wp[q] += eps
        wm[q] -= eps
        lp = UO.Net(wp, bf=False).loss_grad(x, t)[0]
        lm = UO.Net(wm, bf=False).loss_grad(x, t)[0]
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(fd)), (q, fd, g[q])
