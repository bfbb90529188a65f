"""The learned preconditioner on the GPU: the MFMA lane layout, the forward
pass and every parameter gradient against the NumPy restatement (BF16 values
rounded alike, so the two differ only by FP32-vs-float64 summation and the
rare BF16 rounding flip that causes), training lowers the loss, and the
trained network inside the flexible PCG beats the same method without it
(untrained network) and the Jacobi PCG."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import unet_oracle as UO

pytestmark = pytest.mark.gpu


def test_mfma_lane_layout(gpu):
    rng = np.random.default_rng(0)
    A = rng.integers(-8, 8, (16, 16)).astype(np.float32)
    B = rng.integers(-8, 8, (16, 16)).astype(np.float32)
    D = gpu.mfma_check(A, B)
    np.testing.assert_array_equal(D, A @ B)


def sample_input(shape, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros((4,) + tuple(s + 2 for s in shape), dtype=np.float32)
    x[:, 1:-1, 1:-1, 1:-1] = rng.standard_normal((4,) + shape).astype(np.float32)
    x[1:] *= 0.3
    return x


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("shape", [(8, 8, 8), (8, 12, 4)])
def test_forward_matches_restatement(gpu, shape):
    print(f"unet forward {shape}", flush=True)
    net = gpu.UNet(seed=3)
    x = sample_input(shape, 1)
    out = net.forward(x)[1:-1, 1:-1, 1:-1]
    ref = UO.from_internal(UO.Net(net.get_weights(), bf=True).forward(UO.to_internal(x)))
    assert rel(out, ref) < 1e-2, rel(out, ref)


def test_gradients_match_restatement(gpu):
    print("unet gradients", flush=True)
    shape = (8, 8, 8)
    net = gpu.UNet(seed=5)
    x = sample_input(shape, 2)
    t = np.zeros(tuple(s + 2 for s in shape), dtype=np.float32)
    t[1:-1, 1:-1, 1:-1] = np.random.default_rng(3).standard_normal(shape)
    loss, g = net.grad(x, t)
    ref = UO.Net(net.get_weights(), bf=True)
    rloss, rg, _ = ref.loss_grad(UO.to_internal(x), np.transpose(t[1:-1, 1:-1, 1:-1], (2, 1, 0)).astype(np.float64))
    assert abs(loss - rloss) <= 1e-3 * rloss
    off = 0
    for l, (ci, co) in enumerate(UO.LAYERS):
        nw = 27 * ci * co
        gw, rw = g[off:off + nw], rg[off:off + nw]
        gb, rb = g[off + nw:off + nw + co], rg[off + nw:off + nw + co]
        off += nw + co
        assert rel(gw, rw) < 3e-2, (l, rel(gw, rw))
        assert rel(gb, rb) < 3e-2, (l, rel(gb, rb))
    cos = float(g @ rg / (np.linalg.norm(g) * np.linalg.norm(rg)))
    assert cos > 0.999, cos


def test_training_lowers_the_loss(gpu):
    print("unet training", flush=True)
    from moons_amd import learned

    hs = [O.uniform_nodes(0, 1, 16)] * 3
    pairs = learned.random_pairs(hs, ["Neumann"] * 6, 24, seed=4)
    net = gpu.UNet(seed=0)
    hist = net.train([p[0] for p in pairs], [p[1] for p in pairs], epochs=12, batch=4, lr=2e-3, seed=1)
    assert np.all(np.isfinite(hist))
    assert hist[-1] < 0.5 * hist[0], hist


def test_learned_preconditioner_ablation(gpu):
    """32^3 all-Neumann PPE: flexible PCG with the trained network needs far
    fewer iterations than with the untrained network and than Jacobi PCG."""
    print("unet ablation", flush=True)
    from moons_amd import learned

    N = 32
    hs = [O.uniform_nodes(0, 1, N)] * 3
    types = ["Neumann"] * 6
    pairs = learned.random_pairs(hs, types, 48, seed=7)
    trained = gpu.UNet(seed=0)
    hist = trained.train([p[0] for p in pairs], [p[1] for p in pairs], epochs=15, batch=4, lr=1e-3, seed=2)
    print("training loss", hist, flush=True)
    assert np.all(np.isfinite(hist)) and hist[-1] < hist[0]
    untrained = gpu.UNet(seed=0)
    m = gpu.mesh(*hs)
    f = learned.random_rhs((N + 2,) * 3, np.random.default_rng(99))
    b = gpu.SF(m, "CC").upload(f)
    its = {}
    for name, net in (("trained", trained), ("untrained", untrained), ("jacobi", None)):
        x = gpu.SF(m, "CC").fill(0.0).set_BCs(types, None, all_Neumann=True)
        P = gpu.PCG_solver_SF(m, x, gpu.iter_solver_params(3000, 1e-8, 1e-300, 1),
                              prec="UNet" if net is not None else "Lap_SF")
        if net is not None:
            P.set_UNet(net)
        info = P.solve(x, b)
        assert info.res_norm / info.res_norm0 < 1e-8, (name, info)  # Jacobi redo when the learned solve fails
        fell_back = bool(info.exit_loop & 8)
        if name == "trained":
            assert not fell_back
        its[name] = 10**9 if fell_back else info.iterations
    print(its, flush=True)
    assert its["trained"] < its["untrained"] and its["trained"] < its["jacobi"], its
