"""GPU checks of the learned warm start (the U-Net of the learned FCG).

* the MFMA operand layout: an exact integer product through the lane
  mapping every convolution kernel uses;
* the forward pass against a float32 PyTorch (CPU) restatement of the same
  network with the same weights, BF16 rounding emulated where the device
  rounds (weights, input, every activation): relative L2 error < 2e-2;
* the parameter gradient of the training loss against PyTorch autograd of
  the same restatement, per layer: relative L2 error < 8e-2 (activation
  gradients are BF16 on the device);
* training lowers the loss on Poisson problems solved exactly by DCT-II
  (the closed form of the Neumann Laplacian);
* the learned preconditioner inside the cavity solver: steps match the
  oracle as the multigrid path does, whether the prediction is kept or
  rejected.
"""

import numpy as np
import pytest
import scipy.fft as sf

import flursys_amd as fa
import oracle as O
from flursys_amd.unet import UNet, layer_table, params_to_layers, probe_mfma_bf16

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
F = torch.nn.functional


def bf(t):
    return t.to(torch.bfloat16).float()


def torch_unet(weights, x):
    """Float32 restatement of the network (CPU), BF16 rounding emulated."""
    W = [bf(w) for w in weights]

    def conv(t, i, relu=True):
        y = F.conv3d(t, W[i], padding=1)
        return bf(F.relu(y) if relu else y)

    def up(t):
        return F.interpolate(t, scale_factor=2, mode="nearest")

    def pool(t):
        return bf(F.avg_pool3d(t, 2))

    B, n = x.shape[0], round(x.shape[1] ** (1 / 3))
    xt = torch.as_tensor(x, dtype=torch.float32).reshape(B, n, n, n)
    s = torch.sqrt((xt * xt).mean(dim=(1, 2, 3), keepdim=True))
    x0 = torch.zeros(B, 16, n, n, n)
    x0[:, 0] = bf(xt / s)
    a0 = conv(x0, 0); s0 = conv(a0, 1)
    a1 = conv(pool(s0), 2); s1 = conv(a1, 3)
    a2 = conv(pool(s1), 4); s2 = conv(a2, 5)
    a3 = conv(pool(s2), 6); b3 = conv(a3, 7)
    d2 = conv(conv(torch.cat([s2, up(b3)], 1), 8), 9)
    d1 = conv(conv(torch.cat([s1, up(d2)], 1), 10), 11)
    d0 = conv(conv(torch.cat([s0, up(d1)], 1), 12), 13)
    return conv(d0, 14, relu=False), s.reshape(B)


def torch_loss(weights, x, t, diag, gain):
    out, s = torch_unet(weights, x)
    B = x.shape[0]
    o = out[:, 0].reshape(B, -1)
    tt = torch.as_tensor(t, dtype=torch.float32).reshape(B, -1) * (diag / gain / s[:, None])
    e = (o - o.mean(1, keepdim=True)) - (tt - tt.mean(1, keepdim=True))
    return (e * e).mean(1).mean()


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def poisson_pairs(n, batch, seed):
    """Smooth zero-mean right-hand sides and their exact Neumann solutions
    (DCT-II diagonalises the cell-centred Laplacian); h = 1/n."""
    rng = np.random.default_rng(seed)
    k = np.arange(n)
    lam1 = (2.0 - 2.0 * np.cos(np.pi * k / n)) * n * n
    lam = lam1[:, None, None] + lam1[None, :, None] + lam1[None, None, :]
    lam[0, 0, 0] = 1.0
    bs, ps = [], []
    for _ in range(batch):
        bh = rng.standard_normal((n, n, n)) / (1.0 + lam / (4.0 * n * n))
        bh[0, 0, 0] = 0.0
        b = sf.idctn(bh, type=2, norm="ortho")
        p = sf.idctn(bh / lam, type=2, norm="ortho")
        bs.append(b.ravel())
        ps.append(p.ravel())
    return np.array(bs, np.float32), np.array(ps, np.float32), 6.0 * n * n


def typical_gain(x, p, diag):
    """Median of rms(p - mean p) diag / rms(b): makes training targets O(1)."""
    pz = p - p.mean(axis=1, keepdims=True)
    return float(np.median(np.sqrt((pz ** 2).mean(1)) * diag / np.sqrt((x ** 2).mean(1))))


def test_mfma_operand_layout(gpu_ctx):
    rng = np.random.default_rng(0)
    a = rng.integers(-8, 9, (16, 32)).astype(np.float32)
    b = rng.integers(-8, 9, (32, 16)).astype(np.float32)
    d = probe_mfma_bf16(a, b, gpu_ctx)
    assert np.array_equal(d, a @ b)


def test_layer_table_matches_library(gpu_ctx):
    net = UNet(32, 1, seed=3, ctx=gpu_ctx)
    c1, c2, cin, cout, kpad, level, relu, off = layer_table()[-1]
    assert net.num_params == off + cout * kpad + 1  # weights, then the output gain
    w = params_to_layers(net.params())
    assert np.all(w[0][:, 1:] == 0.0)      # input padding channels
    assert np.all(w[-1][1:] == 0.0)        # unused head outputs
    assert np.std(w[1]) > 0.0


@pytest.mark.parametrize("n", [32, 64])
def test_forward_matches_torch(gpu_ctx, n):
    print(f"[gpu] U-Net forward {n}^3", flush=True)
    net = UNet(n, 2, seed=5, ctx=gpu_ctx)
    x, _, _ = poisson_pairs(n, 2, seed=1)
    x += 0.1 * np.random.default_rng(2).standard_normal(x.shape).astype(np.float32)
    raw = net.forward_raw(x)
    assert np.all(raw[..., 1:] == 0.0)
    ref, _ = torch_unet([torch.as_tensor(w) for w in params_to_layers(net.params())], x)
    ref0 = ref[:, 0].reshape(2, -1).numpy()
    assert rel(raw[..., 0], ref0) < 2e-2


def test_infer_is_the_denormalised_zero_mean_output(gpu_ctx):
    n = 32
    net = UNet(n, 1, seed=7, ctx=gpu_ctx)
    x, _, diag = poisson_pairs(n, 1, seed=3)
    raw = net.forward_raw(x)[0, :, 0].astype(np.float64)
    s = np.sqrt(np.mean(x[0].astype(np.float64) ** 2))
    assert net.gain == pytest.approx(0.5 * n * n)   # the default gain
    want = (raw - raw.mean()) * s * net.gain / diag
    got = net.infer(x, diag)[0]
    assert rel(got, want) < 1e-5
    net.set_gain(7.0)
    assert rel(net.infer(x, diag)[0], want * 7.0 / (0.5 * n * n)) < 1e-5
    with pytest.raises(fa.FlursysError):
        net.set_gain(-1.0)


def test_gradients_match_torch(gpu_ctx):
    n = 32
    print("[gpu] U-Net gradients 32^3", flush=True)
    net = UNet(n, 2, seed=11, ctx=gpu_ctx)
    x, p, diag = poisson_pairs(n, 2, seed=4)
    net.set_gain(typical_gain(x, p, diag))
    g, loss = net.grads(x, p, diag)
    ws = [torch.as_tensor(w).clone().requires_grad_(True) for w in params_to_layers(net.params())]
    tl = torch_loss(ws, x, p, diag, net.gain)
    tl.backward()
    assert abs(loss - float(tl)) <= 2e-2 * float(tl)
    for i, (gl, w) in enumerate(zip(params_to_layers(g), ws)):
        tg = w.grad.numpy()
        if i == len(ws) - 1:
            tg, gl = tg[:1], gl[:1]
        if i == 0:
            tg, gl = tg[:, :1], gl[:, :1]
        assert rel(gl, tg) < 8e-2, (i, rel(gl, tg))


def test_training_lowers_the_loss(gpu_ctx):
    n = 32
    print("[gpu] U-Net training 32^3", flush=True)
    net = UNet(n, 4, seed=13, ctx=gpu_ctx)
    x, p, diag = poisson_pairs(n, 4, seed=99)
    net.set_gain(typical_gain(x, p, diag))
    xt, pt, _ = poisson_pairs(n, 2, seed=999)
    pz = pt - pt.mean(axis=1, keepdims=True)
    before = rel(net.infer(xt, diag), pz)
    losses = []
    for step in range(150):
        x, p, diag = poisson_pairs(n, 4, seed=100 + step)
        losses.append(net.train_step(x, p, diag, 2e-3))
    print(f"loss {np.mean(losses[:5]):.4f} -> {np.mean(losses[-10:]):.4f}", flush=True)
    assert np.mean(losses[-10:]) < 0.75 * np.mean(losses[:5])
    # the trained prediction of held-out problems moved towards the solution
    after = rel(net.infer(xt, diag), pz)
    print(f"held-out relative error {before:.3f} -> {after:.3f}", flush=True)
    assert after < 0.9 * before


def test_train_step_validates_arguments(gpu_ctx):
    net = UNet(32, 2, seed=1, ctx=gpu_ctx)
    x, p, diag = poisson_pairs(32, 3, seed=0)
    with pytest.raises(fa.FlursysError):
        net.train_step(x, p, diag, 1e-3)      # batch 3 > max_batch 2
    with pytest.raises(fa.FlursysError):
        net.train_step(x[:1], p[:1], -1.0, 1e-3)
    with pytest.raises(fa.FlursysError):
        UNet(48, 1, ctx=gpu_ctx)              # not a power of two


@pytest.mark.parametrize("grid,netn", [(32, 32), (64, 32)])
def test_learned_steps_match_oracle(gpu_ctx, grid, netn):
    """Whatever the (untrained) network predicts, the solve meets the same
    FP64 stop test and the fields match the oracle: on a grid of the
    network's resolution as a warm start (kept or undone), on a finer grid as
    the coarse solve inside the V-cycle (no separate warm start; a cycle that
    stagnates hands over to the plain FP64 PCG)."""
    print(f"[gpu] learned FCG cavity {grid}^3 with a {netn}^3 U-Net", flush=True)
    kw = dict(reynolds=400.0, dt=5e-4, pressure_tolerance=1e-10, stop_rule=fa.STOP_REL_L2,
              initial_guess=fa.GUESS_ZERO, pressure_max_iters=5000)
    cfg = fa.LidDrivenCavity3DConfig(nx=grid, ny=grid, nz=grid, precision=fa.PRECISION_MIXED,
                                     preconditioner=fa.PRECOND_LEARNED, **kw)
    s = fa.LidDrivenCavity3DSolver(cfg, gpu_ctx)
    net = UNet(netn, 1, seed=21, ctx=gpu_ctx)
    s.set_unet(net)
    o = O.Cavity3D(grid, **{k: v for k, v in kw.items()})
    states = set()
    for _ in range(3):
        rep = s.advance_one_step()
        o.step()
        assert rep.pressure_converged
        states.add(rep.pressure_warm_start)
    if grid == netn:
        assert states <= {1, -1} and states
    else:
        assert states == {0}
    f = s.fields()
    uscale = max(np.abs(o.field(c)).max() for c in ("u", "v", "w"))
    for c in ("u", "v", "w"):
        assert np.abs(f[c] - o.field(c)).max() < 1e-7 * uscale, c
    assert rel(f["p"], o.field("p")) < 1e-7


def test_learned_requires_matching_level(gpu_ctx):
    cfg = fa.LidDrivenCavity3DConfig(nx=32, ny=32, nz=32, precision=fa.PRECISION_MIXED,
                                     preconditioner=fa.PRECOND_LEARNED)
    s = fa.LidDrivenCavity3DSolver(cfg, gpu_ctx)
    with pytest.raises(fa.FlursysError):
        s.set_unet(UNet(64, 1, ctx=gpu_ctx))
    plain = fa.LidDrivenCavity3DSolver(fa.LidDrivenCavity3DConfig(nx=32, ny=32, nz=32), gpu_ctx)
    with pytest.raises(fa.FlursysError):
        plain.set_unet(UNet(32, 1, ctx=gpu_ctx))


def test_coarse_pairs_are_block_means(gpu_ctx):
    cfg = fa.LidDrivenCavity3DConfig(nx=64, ny=64, nz=64, reynolds=400.0,
                                     stop_rule=fa.STOP_REL_L2, pressure_tolerance=1e-8,
                                     initial_guess=fa.GUESS_ZERO, pressure_max_iters=5000,
                                     precision=fa.PRECISION_MIXED,
                                     preconditioner=fa.PRECOND_MULTIGRID)
    s = fa.LidDrivenCavity3DSolver(cfg, gpu_ctx)
    s.advance_one_step()
    b, p, diag = s.coarse_pair(32)
    inter = s.intermediate()["rhs"]
    pf = s.fields()["p"]

    def blocks(a):
        return a.reshape(32, 2, 32, 2, 32, 2).mean(axis=(1, 3, 5)).ravel()

    assert rel(b, blocks(-inter)) < 1e-6
    assert rel(p, blocks(pf)) < 1e-6
    assert diag == pytest.approx(6.0 * 32 * 32)


def test_trained_network_gain_in_fcg(gpu_ctx):
    """What the trained network gains: the learned flexible CG at 128^3 on the
    benchmark's held-out flow (Re 1000) with the committed weights
    (flursys_amd/weights/unet64_cavity.npy) as the V-cycle's 64^3 coarse
    solve, against the same network untrained and against multigrid alone.
    Iterations and ms/step are printed.

    * LEARNED_REPLACE (the network and two sweeps instead of the coarse
      levels): the trained network must converge every step without the
      plain-PCG fallback and need fewer FCG iterations than the untrained
      one -- it does real work there -- but more than multigrid's own
      coarse levels;
    * LEARNED_CORRECTED (the default: the V-cycle corrects the prediction):
      the iteration count is multigrid's, trained or not -- the coarse
      levels already leave no coarse error that limits the outer
      iteration, so the measured gain of the network is zero."""
    import os
    import time

    wpath = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                         "flursys_amd", "weights", "unet64_cavity.npy")
    kw = dict(nx=128, ny=128, nz=128, reynolds=1000.0, dt=5e-4, pressure_tolerance=1e-8,
              stop_rule=fa.STOP_REL_L2, initial_guess=fa.GUESS_ZERO, pressure_max_iters=20000,
              precision=fa.PRECISION_MIXED)
    out = {}
    cases = [("trained", fa.LEARNED_REPLACE), ("untrained", fa.LEARNED_REPLACE),
             ("trained", fa.LEARNED_CORRECTED), ("untrained", fa.LEARNED_CORRECTED),
             ("multigrid", None)]
    for name, mode in cases:
        learned = name != "multigrid"
        cfg = fa.LidDrivenCavity3DConfig(
            preconditioner=fa.PRECOND_LEARNED if learned else fa.PRECOND_MULTIGRID, **kw)
        s = fa.LidDrivenCavity3DSolver(cfg, gpu_ctx)
        if learned:
            s.set_learned_mode(mode)
        net = None
        if learned:
            net = UNet(64, 1, seed=1, ctx=gpu_ctx)
            trained = UNet(64, 1, ctx=gpu_ctx)
            trained.load(wpath)
            if name == "trained":
                net.set_params(trained.params())
            else:  # same architecture and output gain, initial weights
                net.set_gain(trained.gain)
            trained.close()
            s.set_unet(net)
        s.advance_one_step()  # from rest: warm-up
        its, fbs = [], []
        t0 = time.perf_counter()
        for _ in range(4):
            rep = s.advance_one_step()
            assert rep.pressure_converged
            its.append(rep.pressure_iterations)
            fbs.append(rep.pressure_fallback)
        ms = (time.perf_counter() - t0) * 250.0
        tag = name if mode is None else f"{name}/{'replace' if mode else 'corrected'}"
        out[tag] = (its, fbs, ms)
        print(f"[gpu] learned FCG 128^3 Re 1000, {tag}: FCG iterations {its}, "
              f"fallbacks {sum(fbs)}, {ms:.2f} ms/step", flush=True)
        s.close()
        if net is not None:
            net.close()
    its_t, fb_t, _ = out["trained/replace"]
    its_u, fb_u, _ = out["untrained/replace"]
    its_m, fb_m, _ = out["multigrid"]
    assert sum(fb_t) == 0
    assert sum(fb_u) > 0 or sum(its_t) < sum(its_u)
    assert sum(its_t) > sum(its_m)
    its_tc, fb_tc, _ = out["trained/corrected"]
    its_uc, fb_uc, _ = out["untrained/corrected"]
    assert sum(fb_tc) == 0 and sum(fb_uc) == 0
    assert abs(sum(its_tc) - sum(its_m)) <= 2 and abs(sum(its_uc) - sum(its_m)) <= 2
