"""The learned preconditioner on the GPU: the MFMA lane layout, the forward
pass and every parameter gradient against the NumPy restatement (BF16 values
rounded alike, so the two differ only by FP32-vs-float64 summation and the
rare BF16 rounding flip that causes), training lowers the loss, and an
ablation of the trained network against an untrained one, with and without a
V-cycle around it."""
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
    """32^3 all-Neumann PPE, held-out right-hand side.  The trained network
    approximates the exact solution (up to a constant factor) far better than
    the untrained one; as the fine-level initial guess of a two-level V-cycle
    it needs fewer flexible-PCG iterations than the untrained network and no
    more than the cycle without a network (at 32^3 the multigrid already
    removes what the network models).  Alone in the flexible PCG it does not
    reach the tolerance (it models the smooth part of the error only), and
    the Jacobi redo then meets it."""
    print("unet ablation", flush=True)
    from moons_amd import learned

    N = 32
    hs = [O.uniform_nodes(0, 1, N)] * 3
    types = ["Neumann"] * 6
    pairs = learned.random_pairs(hs, types, 48, seed=7)
    trained = gpu.UNet(seed=0)
    hist = trained.train([p[0] for p in pairs], [p[1] for p in pairs], epochs=15, batch=4, lr=1e-3, seed=2)
    print("training loss", hist, flush=True)
    assert np.all(np.isfinite(hist)) and hist[-1] < 0.6 * hist[0]
    untrained = gpu.UNet(seed=0)
    m = gpu.mesh(*hs)
    f = learned.random_rhs((N + 2,) * 3, np.random.default_rng(99))
    e = learned.solve_targets(m, types, [f])[0]
    x_in, t = learned.make_sample(f, e, hs)
    inner = (slice(1, -1),) * 3
    err = {}
    for name, net in (("trained", trained), ("untrained", untrained)):
        o = net.forward(x_in)[inner].astype(np.float64)
        tt = t[inner].astype(np.float64)
        c = (o * tt).sum() / max((o * o).sum(), 1e-300)
        err[name] = float(np.linalg.norm(c * o - tt) / np.linalg.norm(tt))
    print("relative error after scaling", err, flush=True)
    assert err["trained"] < 0.75 * err["untrained"], err
    b = gpu.SF(m, "CC").upload(f)
    # the network as the initial guess of the fine level of a two-level V-cycle
    # (coarse 16^3 level solved by 10 sweeps only): iterations to 1e-8
    its = {}
    for name, net in (("mg2", None), ("mg2+trained", trained), ("mg2+untrained", untrained)):
        x = gpu.SF(m, "CC").fill(0.0).set_BCs(types, None, all_Neumann=True)
        P = gpu.PCG_solver_SF(m, x, gpu.iter_solver_params(500, 1e-8, 1e-300, 1), prec="MG_FP32")
        P.set_MG(2, 2, 2, 10)
        if net is not None:
            P.set_MG_UNet(net, 0)
        info = P.solve(x, b)
        assert info.res_norm / info.res_norm0 < 1e-8, (name, info)
        its[name] = info.iterations
    print("two-level V-cycle iterations", its, flush=True)
    # the trained network is no worse than the cycle without it and better than an untrained one
    assert its["mg2+trained"] <= its["mg2"] and its["mg2+trained"] < its["mg2+untrained"], its
    # standalone in the flexible PCG: the Jacobi fallback returns a converged x
    x = gpu.SF(m, "CC").fill(0.0).set_BCs(types, None, all_Neumann=True)
    P = gpu.PCG_solver_SF(m, x, gpu.iter_solver_params(400, 1e-8, 1e-300, 1), prec="UNet")
    P.set_UNet(trained)
    info = P.solve(x, b)
    print("standalone learned FPCG", info, flush=True)
    assert info.res_norm / info.res_norm0 < 1e-8
