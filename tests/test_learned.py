"""The pressure network and the learned solve.

CPU: the PyTorch statement of the network (tests/unet_ref.py) is pinned --
its convolution follows the documented tap order (checked against a direct
loop), its loss has the closed-form values, and its gradient agrees with a
central finite difference in FP64.

GPU: the library's BF16 matrix-core forward and backward agree with that
statement to BF16 accuracy; capture fills the training set; training lowers
the loss; a learned solve ends at the same residual as the plain solve.
"""
import numpy as np
import pytest
import torch

from tests import unet_ref as R

# ------------------------------------------------------------------ CPU pins


def test_ref_conv_follows_the_documented_tap_order():
    rng = np.random.default_rng(0)
    nb, cin, cout, H, W = 2, 3, 4, 5, 6
    t = rng.standard_normal((nb, cin, H, W))
    w = rng.standard_normal((cout, 9, cin))
    b = rng.standard_normal(cout)
    got = R.conv3x3(torch.tensor(t), torch.tensor(w), torch.tensor(b)).numpy()
    want = np.zeros((nb, cout, H, W))
    for n in range(nb):
        for co in range(cout):
            for h in range(H):
                for x in range(W):
                    acc = b[co]
                    for tap in range(9):
                        dh, dw = tap // 3 - 1, tap % 3 - 1
                        if 0 <= h + dh < H and 0 <= x + dw < W:
                            acc += np.dot(t[n, :, h + dh, x + dw], w[co, tap])
                    want[n, co, h, x] = acc
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_ref_loss_closed_forms():
    rng = np.random.default_rng(1)
    x = torch.tensor(rng.standard_normal((2, 8, 8, 2)))
    x[..., 1] = torch.tensor(rng.random((2, 8, 8)) > 0.3, dtype=torch.float64)
    t = torch.tensor(rng.standard_normal((2, 8, 8)))
    assert float(R.loss(t.clone(), x, t)) == 0.0
    assert abs(float(R.loss(torch.zeros_like(t), x, t)) - 1.0) < 1e-14
    # only masked cells count
    y = t.clone()
    y[x[..., 1] < 0.5] += 100.0
    assert float(R.loss(y, x, t)) == 0.0


def _inputs(nb, H, W, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros((nb, H, W, 2), np.float32)
    x[..., 1] = 1.0
    for n in range(nb):   # a few rectangles of solid cells
        for _ in range(2):
            h0, w0 = rng.integers(0, H - 3), rng.integers(0, W - 3)
            x[n, h0:h0 + 3, w0:w0 + 3, 1] = 0.0
    x[..., 0] = rng.standard_normal((nb, H, W)) * x[..., 1]
    t = (rng.standard_normal((nb, H, W)) * x[..., 1]).astype(np.float32)
    return x, t


def _params(seed, scale_bias=0.1):
    """He-scaled weights, non-zero biases."""
    rng = np.random.default_rng(seed)
    lay, n = R.layout()
    p = np.zeros(n, np.float32)
    for name, off, shape in lay:
        size = int(np.prod(shape))
        if name.endswith(".w") and name.startswith("conv"):
            fan = 9 * shape[2]
            p[off:off + size] = rng.standard_normal(size) * np.sqrt(2.0 / fan)
        elif name == "head.w":
            p[off:off + size] = rng.standard_normal(size) * np.sqrt(1.0 / 32)
        else:
            p[off:off + size] = rng.standard_normal(size) * scale_bias
    return p


def test_ref_gradient_matches_finite_difference():
    x, t = _inputs(1, 8, 8, 2)
    p = _params(3)
    L0, g = R.gradients(p.astype(np.float64), x, t, dtype=torch.float64)
    rng = np.random.default_rng(4)
    d = rng.standard_normal(p.size)
    lay, _ = R.layout()
    # the library keeps conv weights in BF16: differentiate the FP64 function
    # of the rounded weights, i.e. perturb around the rounded point
    pr = R.unpack(p, dtype=torch.float64)
    flat = np.zeros(p.size)
    for name, off, shape in lay:
        flat[off:off + int(np.prod(shape))] = pr[name].numpy().ravel()

    def f(q):
        P = R.unpack(q, dtype=torch.float64, bf16_weights=False)
        xx = R.to_bf16(x).double()
        return float(R.loss(R.forward(P, xx), xx, torch.tensor(t).double()))

    eps = 1e-6
    fd = (f(flat + eps * d) - f(flat - eps * d)) / (2 * eps)
    assert abs(fd - float(g @ d)) <= 1e-6 * max(1.0, abs(fd)), (fd, float(g @ d))


# ------------------------------------------------------------------ GPU


def _gpu_ctx():
    from stroemung_amd.simulation import Context
    return Context(0)


@pytest.mark.gpu
def test_net_layout_and_forward_match_reference():
    from stroemung_amd.learned import PressureNet
    ctx = _gpu_ctx()
    H, W, nb = 32, 24, 3
    net = PressureNet(ctx, H, W, nb, seed=5)
    lay, n = R.layout()
    assert net.nparam == n
    assert [(a, b, c) for a, b, c in net.layout()] == lay
    p = _params(6)
    net.set_params(p)
    np.testing.assert_array_equal(net.get_params(), p)
    x, _ = _inputs(nb, H, W, 7)
    y = net.forward(x)
    P = R.unpack(p)
    with torch.no_grad():
        want = R.forward(P, R.to_bf16(x)).numpy()
        emu = R.forward(P, R.to_bf16(x), bf16_acts=True).numpy()
        emu64 = R.forward(R.unpack(p, dtype=torch.float64), R.to_bf16(x).double(),
                          bf16_acts=True).numpy()
    n = np.linalg.norm
    err, err16 = n(y - want) / n(want), n(y - emu) / n(emu)
    # the noise floor of BF16 storage: the same storage with FP32 or FP64
    # accumulation already differs by this much (rounding flips propagate)
    floor = n(emu - emu64) / n(emu64)
    print("forward relative L2 error: %.3e vs fp32, %.3e vs bf16-stored (floor %.3e)"
          % (err, err16, floor))
    assert err < 3e-2
    assert err16 < 2 * floor + 1e-3
    net.close()
    ctx.close()


@pytest.mark.gpu
def test_net_gradients_match_reference():
    from stroemung_amd.learned import PressureNet
    ctx = _gpu_ctx()
    H, W, nb = 32, 32, 2
    net = PressureNet(ctx, H, W, nb, seed=8)
    p = _params(9)
    net.set_params(p)
    x, t = _inputs(nb, H, W, 10)
    loss, g = net.gradients(x, t)
    Lr, gr = R.gradients(p, x, t)
    print("loss gpu %.6f ref fp32 %.6f" % (loss, Lr))
    assert abs(loss - Lr) <= 2e-2 * abs(Lr)
    for name, off, shape in R.layout()[0]:
        size = int(np.prod(shape))
        cos, rel = _cmp(g[off:off + size], gr[off:off + size])
        print("%-10s cos %.5f rel %.3e" % (name, cos, rel))
        # BF16 storage of ~20 activations and of their gradients: ReLU and
        # pooling decisions flip and the flips propagate, most at the 8x8
        # bottleneck (conv6.w: cos 0.9938, rel 0.112 -- the fixed-order kernels
        # give the same figures on every run); the head, after the last BF16
        # rounding, is within 1.2e-3.  The exact-arithmetic comparison below
        # holds the kernels to 1e-3.
        assert cos > 0.99 and rel < 0.13, name
        if name.startswith("head"):
            assert rel < 5e-3, name
    # the padding of the first layer's input channels gets no gradient
    w0 = g[0:32 * 9 * 32].reshape(32, 9, 32)
    assert np.all(w0[:, :, 2:] == 0.0)
    net.close()
    ctx.close()


def _cmp(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-300))
    return cos, float(np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-300))


def _sparse_params(seed):
    """Each output channel reads one (tap, input channel) with weight +-1 and
    an integer bias; the head has weights in {0, +-1, 1/2}.  With integer
    inputs every activation is a small integer, exact in BF16, so the forward
    pass has no rounding at all and must match bit for bit."""
    rng = np.random.default_rng(seed)
    lay, n = R.layout()
    p = np.zeros(n, np.float32)
    real_cin = {0: 2}
    for name, off, shape in lay:
        size = int(np.prod(shape))
        if name.startswith("conv") and name.endswith(".w"):
            k = int(name[4:-2])
            cout, _, cin = shape
            w = np.zeros(shape, np.float32)
            for co in range(cout):
                w[co, rng.integers(0, 9), rng.integers(0, real_cin.get(k, cin))] = \
                    1.0 if rng.random() < 0.8 else -1.0
            p[off:off + size] = w.ravel()
        elif name.startswith("conv"):
            p[off:off + size] = rng.integers(-1, 2, size)
        elif name == "head.w":
            p[off:off + size] = rng.choice([0.0, 1.0, -1.0, 0.5], size)
        else:
            p[off:off + size] = 0.5
    return p


@pytest.mark.gpu
@pytest.mark.parametrize("H,W", [(32, 40), (32, 48), (64, 64)])
def test_net_exact_arithmetic_forward_is_bit_identical_and_gradients_tight(H, W):
    """Integer activations: no rounding in the forward pass, so any index
    mistake (tap, channel, pooling window, upsampling, concatenation order,
    head) shows as a difference; the backward pass then rounds only its
    gradients to BF16, at the places the emulated reference does."""
    from stroemung_amd.learned import PressureNet
    ctx = _gpu_ctx()
    nb = 2
    net = PressureNet(ctx, H, W, nb, seed=1)
    p = _sparse_params(13)
    net.set_params(p)
    rng = np.random.default_rng(14)
    x, _ = _inputs(nb, H, W, 15)
    x[..., 0] = rng.integers(-4, 5, (nb, H, W)) * x[..., 1]
    y = net.forward(x)
    with torch.no_grad():
        want = R.forward(R.unpack(p), torch.tensor(x)).numpy()
    assert np.count_nonzero(want) > want.size // 4
    np.testing.assert_array_equal(y, want)
    t = (rng.integers(-3, 4, (nb, H, W)) * x[..., 1]).astype(np.float32)
    loss, g = net.gradients(x, t)
    L16, g16 = R.gradients(p, x, t, bf16_acts=True)
    assert abs(loss - L16) <= 1e-5 * abs(L16)
    for name, off, shape in R.layout()[0]:
        size = int(np.prod(shape))
        cos, rel = _cmp(g[off:off + size], g16[off:off + size])
        print("%-10s cos %.7f rel %.3e" % (name, cos, rel))
        assert rel < 1e-3, name


def _cavity_context(count, n, seed, eps=1e-6):
    from stroemung_amd import workloads
    from stroemung_amd.simulation import Context
    ctx = Context(0)
    grids = workloads.random_cavities(count, n=n, seed=seed, eps=eps)
    for ct, vel, prm in grids:
        ctx.add_grid(prm, ct, vel)
    ctx.construct()
    return ctx, grids


def _ghosts(ct, p):
    """p with the boundary ghosts refreshed by the oracle's rule."""
    from oracle import pyoracle as O
    sim = O.Sim(ct, delx=1.0, dely=1.0, delt=0.0, gamma=0.0, reynolds=1.0,
                sor_absolute_epsilon=1.0, max_iterations=1, omega=1.0, p=p)
    assert sim.call("rebuild_boundary_list") == 0
    sim.call("copy_pressure_to_boundaries")
    return sim.p.copy()


def _restrict(a, f):
    n0, n1 = a.shape
    return a.reshape(n0 // f, f, n1 // f, f).sum(axis=(1, 3)) / (f * f)


@pytest.mark.gpu
@pytest.mark.parametrize("f", [1, 4, 16])
def test_capture_pairs_follow_the_restriction(f):
    """The stored pair of the first capture point, recomputed from the fields:
    residual from ghost-refreshed p, restriction, scale, fluid fraction and
    the target from the converged p."""
    from stroemung_amd.learned import PressureNet
    n, count = (128 if f == 16 else 64), 2       # frames are at least 8 x 8
    ctx, grids = _cavity_context(count, n, seed=21)
    for _ in range(3):
        ctx.tick()
    p0 = [ctx.download(c, ("p",))["p"] for c in range(count)]
    net = PressureNet(ctx, n // f, n // f, 4, seed=1)
    net.reserve(8)
    ctx.set_capture(net, points=2, sweeps=8, coarsen=f)
    ctx.tick()
    ctx.set_capture(None)
    for c, (ct, _, prm) in enumerate(grids):
        st = ctx.state(c)
        fld = ctx.download(c, ("p", "rhs"))
        pg = _ghosts(ct, p0[c])
        fl = (ct == 0)[1:-1, 1:-1]
        P = pg
        pc = P[1:-1, 1:-1]
        rdx2, rdy2 = 1.0 / prm.delx ** 2, 1.0 / prm.dely ** 2
        lap = (((P[2:, 1:-1] - pc) - (pc - P[:-2, 1:-1])) * rdx2
               + ((P[1:-1, 2:] - pc) - (pc - P[1:-1, :-2])) * rdy2)
        r = np.where(fl, lap - (fld["rhs"][1:-1, 1:-1] - st["rhs_mean"]), 0.0)
        rc = _restrict(r, f)
        frac = _restrict(fl.astype(float), f)
        scale = np.sqrt(np.count_nonzero(frac) / np.sum(rc * rc))
        x, t = net.sample(c)
        np.testing.assert_array_equal(x[..., 1], R.to_bf16(frac).numpy())
        want_x = -rc * scale
        assert np.max(np.abs(x[..., 0] - want_x)) <= 1e-2 * np.max(np.abs(want_x)) + 1e-6
        d = np.where(fl, fld["p"][1:-1, 1:-1] - p0[c][1:-1, 1:-1], 0.0)
        d = np.where(fl, d - d[fl].mean(), 0.0)
        want_t = _restrict(d, f) * scale / (f * f * prm.delx * prm.dely)
        np.testing.assert_allclose(t, want_t, rtol=1e-4, atol=1e-5 * np.max(np.abs(want_t)))
    net.close()
    ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("f,step", [(1, "network"), (4, "network"), (4, "energy"),
                                    (4, "momentum"), (4, "momentum2"), (16, "momentum")])
def test_capture_train_and_learned_solve_reach_the_same_residual(f, step):
    from stroemung_amd.learned import PressureNet
    n, count = (128 if f == 16 else 64), 4       # frames are at least 8 x 8
    ctx, _ = _cavity_context(count, n, seed=11)
    net = PressureNet(ctx, n // f, n // f, 32, seed=12)
    net.reserve(64)
    ctx.set_capture(net, points=2, sweeps=16, coarsen=f)
    for _ in range(6):
        its, _ = ctx.tick()
        for c in range(count):
            s = ctx.state(c)
            assert s["residual_squared"] <= s["tolerance_squared"] and not s["exhausted"]
    ctx.set_capture(None)
    assert net.dataset_size() == 6 * 2 * count
    first = net.train(20, 16, lr=1e-3, seed=1)
    last = net.train(200, 16, lr=1e-3, seed=2)
    print("training loss %.4f -> %.4f" % (first, last))
    assert last < first
    # learned solve: same stopping test, so the same residual bound
    ctx.set_learned(net, cycles=2, sweeps=16, step=step, coarsen=f)
    for _ in range(3):
        its, _ = ctx.tick()
        for c in range(count):
            s = ctx.state(c)
            assert s["residual_squared"] <= s["tolerance_squared"] and not s["exhausted"]
            assert s["sor_iterations"] == its[c]
    st = net.stats()
    assert st["forward_calls"] >= 3 and st["forward_ms"] > 0
    # capture inside learned solves (pairs from the learned solver's own states)
    before = net.dataset_size()
    ctx.set_capture(net, points=2, sweeps=16, coarsen=f)
    its, _ = ctx.tick()
    for c in range(count):
        s = ctx.state(c)
        assert s["residual_squared"] <= s["tolerance_squared"] and not s["exhausted"]
    assert net.dataset_size() == min(64, before + 2 * count)
    ctx.set_capture(None)
    ctx.set_learned(None)
    net.close()
    ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("step", ["energy", "momentum", "momentum2"])
def test_graph_replayed_cycles_are_bit_identical_to_launches(step):
    """Learned cycles replayed from a captured hipGraph give the same fields,
    bit for bit, as the same cycles launched one by one -- over several ticks,
    each replaying the graph several times between host synchronisations."""
    from stroemung_amd.learned import PressureNet
    n, count, f = 64, 4, 4
    runs = []
    for graphs in (False, True):
        ctx, _ = _cavity_context(count, n, seed=31)
        net = PressureNet(ctx, n // f, n // f, count, seed=32)
        for _ in range(2):
            ctx.tick()
        ctx.set_learned(net, cycles=6, sweeps=8, step=step, coarsen=f)
        ctx.set_graphs(graphs)
        its = [ctx.tick()[0] for _ in range(4)]
        fields = [ctx.download(c, ("u", "v", "p")) for c in range(count)]
        for c in range(count):
            s = ctx.state(c)
            assert s["residual_squared"] <= s["tolerance_squared"] and not s["exhausted"]
        runs.append((its, fields))
        ctx.set_learned(None)
        net.close()
        ctx.close()
    assert runs[0][0] == runs[1][0]
    for a, b in zip(runs[0][1], runs[1][1]):
        for k in ("u", "v", "p"):
            np.testing.assert_array_equal(a[k], b[k])


@pytest.mark.gpu
def test_trained_network_saves_sweeps_on_unseen_layouts():
    """What the learned part gains: a network trained (on the GPU, with the
    library's kernels) on pairs from 64x64 cavities of one seed needs far
    fewer red-black sweeps than the plain solver on cavities of another
    seed, from the same state and to the same residual."""
    from stroemung_amd import learned, workloads
    train = workloads.random_cavities(16, n=64, seed=41, eps=1e-6)
    params, info = learned.train_net(train, coarsen=4, capture_ticks=24, points=4,
                                     capture_sweeps=16, steps=600, batch=32, dagger_rounds=1,
                                     dagger_ticks=8, dagger_steps=300, cycles=12, sweeps=16)
    test = workloads.random_cavities(8, n=64, seed=42, eps=1e-6)
    res = learned.compare(test, params, coarsen=4, cycles=12, sweeps=16, warm=6, ticks=4)
    print("plain %.1f sweeps/solve, learned %.1f" % (res["plain"]["sweeps_per_solve"],
                                                    res["learned"]["sweeps_per_solve"]))
    assert res["plain"]["residual_bound_met"] and res["learned"]["residual_bound_met"]
    assert res["learned"]["sweeps_per_solve"] < 0.5 * res["plain"]["sweeps_per_solve"]
    # the network is fully convolutional: the same weights on larger grids and
    # frames (128 x 128 cavities, 32 x 32 frames) still cut the sweeps
    big = workloads.random_cavities(2, n=128, seed=43, eps=1e-6)
    res = learned.compare(big, params, coarsen=4, cycles=24, sweeps=16, warm=3, ticks=2,
                          step="momentum")
    print("128x128: plain %.1f sweeps/solve, learned %.1f" % (
        res["plain"]["sweeps_per_solve"], res["learned"]["sweeps_per_solve"]))
    assert res["plain"]["residual_bound_met"] and res["learned"]["residual_bound_met"]
    assert res["learned"]["sweeps_per_solve"] < 0.5 * res["plain"]["sweeps_per_solve"]


@pytest.mark.gpu
def test_trajectory_runs_learned_ticks_with_plain_checkpoints():
    """The long-run helper: learned ticks from rest, each at its residual
    bound, with plain ticks timed on the same state at the checkpoints."""
    from stroemung_amd import learned, workloads
    grids = workloads.fixed_channels(4, seed=5, imax=64, jmax=64, eps=1e-6)
    ctx = learned.make_context(grids)
    net = learned.PressureNet(ctx, 16, 16, 4, seed=3)
    params = net.get_params()
    net.close()
    ctx.close()
    res = learned.trajectory(grids, params, coarsen=4, cycles=4, sweeps=8, ticks=12,
                             checkpoints=(5, 12))
    assert res["learned"]["residual_bound_met"]
    assert [m["tick"] for m in res["checkpoints"]] == [5, 12]
    for m in res["checkpoints"]:
        assert m["plain_ms_per_tick"] > 0 and m["learned_ms_per_tick"] > 0
        assert m["plain_sweeps_per_solve"] > 0


@pytest.mark.gpu
def test_small_batch_forward_is_deterministic():
    """One image of 64 x 64 splits the deep layers' input channels over
    several workgroups; their partial sums are added in a fixed order, so
    repeated forwards give the same bits."""
    from stroemung_amd.learned import PressureNet
    ctx, _ = _cavity_context(1, 64, seed=51)
    net = PressureNet(ctx, 64, 64, 1, seed=52)
    x = np.random.default_rng(53).standard_normal((1, 64, 64, 2)).astype(np.float32)
    x[..., 1] = 1.0
    ys = [net.forward(x) for _ in range(4)]
    for y in ys[1:]:
        np.testing.assert_array_equal(y, ys[0])
    assert np.isfinite(ys[0]).all() and np.abs(ys[0]).max() > 0
    net.close()
    ctx.close()


@pytest.mark.gpu
def test_gradients_and_training_are_bitwise_reproducible():
    """The weight, bias and head gradients and the loss are sums of per-block
    partials added in a fixed order.  At 8 images of 64 x 64 the first
    level's weight gradient has eight pixel chunks and the bias and head
    reductions several blocks each, so repeated gradients and two trainings
    from the same start and the same captured set give the same bits."""
    from stroemung_amd.learned import PressureNet
    ctx = _gpu_ctx()
    H, W, nb = 64, 64, 8
    net = PressureNet(ctx, H, W, nb, seed=61)
    x, t = _inputs(nb, H, W, 62)
    runs = [net.gradients(x, t) for _ in range(3)]
    for loss, g in runs[1:]:
        assert loss == runs[0][0]
        np.testing.assert_array_equal(g, runs[0][1])
    assert np.isfinite(runs[0][1]).all() and np.abs(runs[0][1]).max() > 0
    net.close()
    ctx.close()

    trained = []
    for _ in range(2):
        ctx, _ = _cavity_context(2, 64, seed=63)
        for _ in range(2):
            ctx.tick()
        net = PressureNet(ctx, 16, 16, 4, seed=64)
        net.reserve(16)
        ctx.set_capture(net, points=4, sweeps=8, coarsen=4)
        for _ in range(2):
            ctx.tick()
        ctx.set_capture(None)
        loss = net.train(24, 4, lr=1e-3, seed=65)
        trained.append((net.dataset_size(), loss, net.get_params()))
        net.close()
        ctx.close()
    (n0, l0, p0), (n1, l1, p1) = trained
    assert n0 == n1 > 4
    assert l0 == l1 and np.isfinite(l0)
    np.testing.assert_array_equal(p0, p1)


def _masked_residual(ct, p, rhs, mean, dx, dy):
    """The oracle's sum of r^2 over interior fluid cells (ghosts refreshed)."""
    from oracle import pyoracle as O
    sim = O.Sim(ct, delx=dx, dely=dy, delt=0.0, gamma=0.0, reynolds=1.0,
                sor_absolute_epsilon=1.0, max_iterations=1, omega=1.8, p=p)
    sim.rhs[...] = rhs
    sim.set(rhs_mean=mean)
    assert sim.call("rebuild_boundary_list") == 0
    return O.lib().orc_masked_residual_sum(sim.s, None)


@pytest.mark.gpu
@pytest.mark.parametrize("itermax,cycles,sweeps", [(10, 3, 16), (36, 3, 16), (37, 8, 6)])
def test_learned_solve_stops_at_itermax(itermax, cycles, sweeps):
    """A learned solve never sweeps past max_iterations: a phase whose budget
    would pass it is cut to the sweeps left.  With 4 sweeps per launch the
    cut phases end after an odd number of sweep-doing launches while the
    phase queued an even number (36 = 16 + 16 + 4; 37 = 6 * 6 + 1), so p's
    latest buffer must come from the launch that stopped the channel.  Each
    channel reports exactly itermax sweeps, exhausted, and the residual it
    reports is that of the p it leaves."""
    from stroemung_amd import workloads
    from stroemung_amd.learned import PressureNet
    from stroemung_amd.simulation import Context
    grids = workloads.random_cavities(3, n=64, seed=71, eps=1e-13, check_every=4,
                                      itermax=itermax)
    for graphs in (False, True):
        ctx = Context(0)
        for ct, vel, prm in grids:
            ctx.add_grid(prm, ct, vel)
        ctx.construct()
        net = PressureNet(ctx, 16, 16, 3, seed=72)
        ctx.set_learned(net, cycles=cycles, sweeps=sweeps, step="momentum", coarsen=4)
        ctx.set_graphs(graphs)
        for tick in range(2):
            its, _ = ctx.tick()
            for c, (ct, _, prm) in enumerate(grids):
                st = ctx.state(c)
                assert its[c] == itermax and st["sor_iterations"] == itermax, (graphs, tick, c)
                assert st["exhausted"], (graphs, tick, c)
                fld = ctx.download(c, ("p", "rhs"))
                r2 = _masked_residual(ct, fld["p"], fld["rhs"], st["rhs_mean"], prm.delx,
                                      prm.dely)
                assert st["residual_squared"] == pytest.approx(r2, rel=1e-10), (graphs, tick, c)
        ctx.set_learned(None)
        net.close()
        ctx.close()


@pytest.mark.gpu
def test_capture_phases_stop_at_itermax():
    """Capture phases run 8, 16, 32, ... sweeps; with itermax 20 the second is
    cut to 12 and the solve ends exhausted after 20 sweeps -- the same p, u
    and v, bit for bit, as the plain solve stopped at 20 sweeps."""
    from stroemung_amd import workloads
    from stroemung_amd.learned import PressureNet
    from stroemung_amd.simulation import Context
    grids = workloads.random_cavities(2, n=64, seed=73, eps=1e-13, check_every=4, itermax=20)
    runs = []
    for capture in (False, True):
        ctx = Context(0)
        for ct, vel, prm in grids:
            ctx.add_grid(prm, ct, vel)
        ctx.construct()
        net = PressureNet(ctx, 16, 16, 2, seed=74)
        net.reserve(16)
        if capture:
            ctx.set_capture(net, points=4, sweeps=8, coarsen=4)
        its = [ctx.tick()[0] for _ in range(2)]
        for c in range(len(grids)):
            assert ctx.state(c)["exhausted"]
        runs.append((its, [ctx.download(c, ("u", "v", "p")) for c in range(len(grids))]))
        ctx.set_capture(None)
        net.close()
        ctx.close()
    assert runs[0][0] == runs[1][0] == [[20, 20], [20, 20]]
    for a, b in zip(runs[0][1], runs[1][1]):
        for k in ("u", "v", "p"):
            np.testing.assert_array_equal(a[k].view(np.uint64), b[k].view(np.uint64))


def _free_block(ct, rng, size=2, margin=2):
    """Origin of a size x size block with only fluid within `margin` cells."""
    nx, ny = ct.shape
    for _ in range(1000):
        x = int(rng.integers(margin + 1, nx - size - margin - 1))
        y = int(rng.integers(margin + 1, ny - size - margin - 1))
        if (ct[x - margin:x + size + margin, y - margin:y + size + margin] == 0).all():
            return x, y
    raise AssertionError("no free block")


@pytest.mark.gpu
def test_graph_replay_after_grid_edits_and_net_swaps():
    """Captured learned cycles hold the boundary list, tile lists and net
    buffers by value; after a grid edit (an obstacle added: a longer list,
    then removed: a shorter one) and after the net is destroyed and another
    of the same shape set, the replayed cycles must equal the same sequence
    launched one by one, bit for bit."""
    from stroemung_amd import workloads
    from stroemung_amd.learned import PressureNet
    from stroemung_amd.simulation import Context
    n, count, f = 64, 3, 4
    grids = workloads.random_cavities(count, n=n, seed=81, eps=1e-6)
    runs = []
    for graphs in (False, True):
        rng = np.random.default_rng(82)
        ctx = Context(0)
        for ct, vel, prm in grids:
            ctx.add_grid(prm, ct, vel)
        ctx.construct()
        net = PressureNet(ctx, n // f, n // f, count, seed=83)
        ctx.set_graphs(graphs)
        ctx.set_learned(net, cycles=6, sweeps=8, step="momentum", coarsen=f)
        log = []
        log.append([ctx.tick()[0] for _ in range(2)])
        ct0, vel0, _ = grids[1]
        x, y = _free_block(ct0, rng)
        edited = ct0.copy()
        edited[x:x + 2, y:y + 2] = 1
        ctx.set_cells(1, edited, vel0)                  # more boundary cells
        log.append([ctx.tick()[0] for _ in range(2)])
        ctx.set_cells(1, ct0, vel0)                     # fewer again
        log.append([ctx.tick()[0] for _ in range(2)])
        net.close()                                     # the context stops using it
        log.append([ctx.tick()[0] for _ in range(1)])
        net = PressureNet(ctx, n // f, n // f, count, seed=84)
        ctx.set_learned(net, cycles=6, sweeps=8, step="momentum", coarsen=f)
        log.append([ctx.tick()[0] for _ in range(2)])
        for c in range(count):
            s = ctx.state(c)
            assert s["residual_squared"] <= s["tolerance_squared"] and not s["exhausted"]
        runs.append((log, [ctx.download(c, ("u", "v", "p")) for c in range(count)]))
        ctx.set_learned(None)
        net.close()
        ctx.close()
    assert runs[0][0] == runs[1][0]
    for a, b in zip(runs[0][1], runs[1][1]):
        for k in ("u", "v", "p"):
            np.testing.assert_array_equal(a[k].view(np.uint64), b[k].view(np.uint64))
