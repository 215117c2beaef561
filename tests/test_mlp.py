"""Warm-start network: the numpy oracle (oracle/mlp.py) pinned by central
finite differences on the CPU, then the device kernels (matrix-core GEMMs,
tanh backward, column sums, loss, Adam) against the oracle on the GPU.
FP32 device arithmetic: bound 1e-5 relative to the largest entry."""
import numpy as np
import pytest

from oracle import mlp as om


def _rand_params(widths, seed):
    rng = np.random.default_rng(seed)
    n = sum(widths[l] * widths[l + 1] + widths[l + 1] for l in range(len(widths) - 1))
    return rng.normal(scale=0.5, size=n)


def test_oracle_gradient_matches_finite_differences():
    widths = [3, 5, 4, 6]
    p = _rand_params(widths, 0)
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, size=(7, 3))
    t = rng.normal(size=(7, 6))
    _, g = om.gradient(widths, p, x, t)
    h = 1e-6
    fd = np.empty_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = h
        fd[i] = (om.loss(widths, p + e, x, t) - om.loss(widths, p - e, x, t)) / (2 * h)
    assert np.max(np.abs(g - fd)) <= 1e-8 * max(1.0, np.max(np.abs(fd)))


def test_oracle_linear_network_is_least_squares():
    """One linear layer: the loss gradient vanishes at the least-squares fit."""
    rng = np.random.default_rng(2)
    x = rng.normal(size=(40, 3))
    t = rng.normal(size=(40, 2))
    X1 = np.hstack([x, np.ones((40, 1))])
    sol = np.linalg.lstsq(X1, t, rcond=None)[0]
    p = np.concatenate([sol[:3].ravel(), sol[3]])
    _, g = om.gradient([3, 2], p, x, t)
    assert np.max(np.abs(g)) < 1e-12


gpu = pytest.mark.gpu


def _net(widths, seed=3):
    from iceicle_amd import engine
    return engine.MLP(widths, seed=seed)


@gpu
def test_forward_and_gradient_match_oracle():
    torch = pytest.importorskip("torch")
    widths = [3, 70, 33, 130]  # no dimension a multiple of the 64 x 64 tiles
    net = _net(widths)
    p = net.get_params()
    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, (77, 3)).astype(np.float32)
    t = rng.normal(size=(77, 130)).astype(np.float32)
    xd, td = torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda()
    y = net.forward(xd).cpu().numpy()
    yr = om.forward(widths, p, x)
    ferr = np.abs(y - yr).max() / np.abs(yr).max()
    loss, g = net.gradient(xd, td)
    lr, gr = om.gradient(widths, p, x, t)
    gerr = np.abs(g.cpu().numpy() - gr).max() / np.abs(gr).max()
    print(f"forward rel err {ferr:.2e}, gradient rel err {gerr:.2e}, loss {loss} vs {lr}", flush=True)
    assert ferr <= 1e-5 and gerr <= 1e-5
    assert abs(loss - lr) <= 1e-5 * lr
    # deterministic reductions: a second evaluation is bit-identical
    loss2, g2 = net.gradient(xd, td)
    assert loss2 == loss and torch.equal(g, g2)


@gpu
def test_wide_output_layer_matches_oracle():
    """The warm-start shape 3 -> 256 -> 256 -> 1024 at a ragged batch."""
    torch = pytest.importorskip("torch")
    widths = [3, 256, 256, 1024]
    net = _net(widths, seed=7)
    p = net.get_params()
    rng = np.random.default_rng(4)
    x = rng.uniform(-1, 1, (100, 3)).astype(np.float32)
    t = rng.normal(size=(100, 1024)).astype(np.float32)
    _, g = net.gradient(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda())
    _, gr = om.gradient(widths, p, x, t)
    err = np.abs(g.cpu().numpy() - gr).max() / np.abs(gr).max()
    print(f"gradient rel err {err:.2e}", flush=True)
    assert err <= 1e-5


@gpu
def test_training_fits_a_smooth_map():
    torch = pytest.importorskip("torch")
    from iceicle_amd import abi
    widths = [3, 64, 64, 40]
    net = _net(widths, seed=1)
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, (512, 3)).astype(np.float32)
    s = np.linspace(0, 1, 40)
    t = (np.sin(np.pi * (x[:, :1] + s)) * x[:, 1:2] + 0.5 * x[:, 2:3] * s).astype(np.float32)
    xd, td = torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda()
    l0, _ = net.gradient(xd, td)
    hist = net.train(xd, td, abi.mlp_train_opts(batch=64, epochs=150, lr=3e-3, lr_decay=0.98, seed=2))
    l1, _ = net.gradient(xd, td)
    print(f"loss {l0:.3e} -> {l1:.3e} (epoch means {hist[0]:.3e} .. {hist[-1]:.3e})", flush=True)
    assert l1 < 0.02 * l0


@gpu
def test_parameter_round_trip_and_errors():
    torch = pytest.importorskip("torch")
    from iceicle_amd import engine
    net = _net([2, 3, 1])
    p = np.arange(net.nparams, dtype=np.float32) * 0.01
    net.set_params(p)
    assert np.array_equal(net.get_params(), p)
    with pytest.raises(engine.IceError):
        net.forward(torch.zeros((4, 3), dtype=torch.float32, device="cuda"))
    with pytest.raises(engine.IceError):
        net.forward(torch.zeros((4, 2), dtype=torch.float64, device="cuda"))


def test_warm_start_features_and_sampling():
    from iceicle_amd import warmstart
    p = warmstart.sample_params(500, seed=9)
    assert np.array_equal(p, warmstart.sample_params(500, seed=9))
    assert p[:, 0].min() >= 1e-3 and p[:, 0].max() <= 1e-2
    assert p[:, 1:].min() >= 0.5 and p[:, 1:].max() <= 2.0
    f = warmstart.features(p)
    assert f.dtype == np.float32 and f.min() >= -1.0 - 1e-6 and f.max() <= 1.0 + 1e-6
    corners = warmstart.features([[1e-3, 0.5, 2.0], [1e-2, 2.0, 0.5]])
    assert np.allclose(corners, [[-1, -1, 1], [1, 1, -1]], atol=1e-6)


def test_batch_instance_sampling():
    from iceicle_amd import batch
    insts = batch.sample_instances(512, seed=3)
    again = batch.sample_instances(512, seed=3)
    assert all(a["N"] == b["N"] and a["p"] == b["p"] and a["nu"] == b["nu"] for a, b in zip(insts, again))
    assert min(i["N"] for i in insts) >= 16 and max(i["N"] for i in insts) <= 96
    assert {i["p"] for i in insts} == {1, 2, 3, 4}
    assert all(1e-3 <= i["nu"] <= 1e-1 for i in insts)
    # the initial condition is the stated sum of sinusoids
    x, y = np.array([0.1, 0.3]), np.array([0.2, 0.7])
    amp = np.array([1.0, 0.0, 0.0, 0.0])
    assert np.allclose(batch.initial_condition(x, y, amp), np.sin(2 * np.pi * x))
