"""GPU checks of the learned denoiser and of the auxiliary buffers it reads.

The BF16 U-Net's forward pass and its gradient are compared with a float32
PyTorch restatement of the same network on the CPU that rounds weights and
stored activations to bfloat16 at the same places the kernels do (products
are exact in fp32 either way; only the summation order differs), with the
tolerances written in each test. Training must reduce the loss on a fixed
set. The auxiliary buffers must not change the image, and per pixel the
transmittance plus the channel fractions is 1 (every sample either misses
or hits exactly one channel). The per-pair loss weights act on the gradient
only, and the inputs read during one accumulation carry its noise level."""
import numpy as np
import pytest

from mdwtracker_amd import scenes

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

LAYERS = [(None, 32, True), (32, 32, True), (32, 64, True), (64, 64, True), (64, 96, True),
          (96, 96, True), (160, 64, True), (64, 64, True), (96, 32, True), (32, 32, True),
          (32, 49, False)]


@pytest.fixture(scope="module")
def dnmod():
    from mdwtracker_amd import denoise
    return denoise


def _bf(t):
    return t.to(torch.bfloat16).to(torch.float32)


def _layers(channels):
    return [(7 + channels if cin is None else cin, cout, relu) for cin, cout, relu in LAYERS]


def _unpack(p, channels):
    out, o = [], 0
    for cin, cout, _ in _layers(channels):
        w = p[o:o + 9 * cout * cin].reshape(9, cout, cin)
        o += 9 * cout * cin
        b = p[o:o + cout]
        o += cout
        out.append((w, b))
    assert o == p.size
    return out


def _torch_forward(params, x9, channels=4):
    """float32 restatement of the U-Net and its kernel-predicting output with
    the kernels' bf16 rounding."""
    F = torch.nn.functional
    layers = []
    for (w, b), (cin, cout, relu) in zip(_unpack(params, channels), _layers(channels)):
        wt = torch.tensor(w).reshape(3, 3, cout, cin).permute(2, 3, 0, 1).contiguous()
        layers.append((_bf(wt).requires_grad_(True), torch.tensor(b).requires_grad_(True), relu))

    def conv(i, x, store_bf16=True):
        w, b, relu = layers[i]
        y = F.conv2d(x, w, b, padding=1)
        if relu:
            y = F.relu(y)
        return _bf(y) if store_bf16 else y

    x = torch.tensor(x9).permute(2, 0, 1)[None]
    H, W = x9.shape[:2]
    x0 = _bf(x)
    a1 = conv(1, conv(0, x0))
    p1 = _bf(F.avg_pool2d(a1, 2))
    a3 = conv(3, conv(2, p1))
    p2 = _bf(F.avg_pool2d(a3, 2))
    a5 = conv(5, conv(4, p2))
    u2 = torch.cat([F.interpolate(a5, scale_factor=2, mode="nearest"), a3], 1)
    a7 = conv(7, conv(6, u2))
    u1 = torch.cat([F.interpolate(a7, scale_factor=2, mode="nearest"), a1], 1)
    a9 = conv(9, conv(8, u1))
    logits = conv(10, a9, store_bf16=False)
    # 7x7 kernel prediction over the clamped neighbourhood of the input colour
    w = torch.softmax(logits, 1)  # [1][49][H][W]
    rgb = F.pad(x[:, :3], (3, 3, 3, 3), mode="replicate")
    patches = F.unfold(rgb, 7).reshape(1, 3, 49, H, W)
    pred = (patches * w[:, None]).sum(2)
    return pred[0].permute(1, 2, 0), layers


def _rand_input(rng, h, w, channels=4):
    x = np.zeros((h, w, 7 + channels), np.float32)
    x[..., :3] = rng.uniform(0, 1.2, (h, w, 3))
    x[..., 3] = rng.uniform(0, 1, (h, w))
    frac = rng.dirichlet(np.ones(channels + 1), size=(h, w)).astype(np.float32)
    x[..., 4:5 + channels] = frac
    x[..., 5 + channels] = rng.uniform(0, 1, (h, w))
    x[..., 6 + channels] = rng.uniform(0, 1)
    return x


@pytest.mark.parametrize("channels,h,w", [(4, 32, 48), (10, 32, 48), (4, 12, 20), (4, 68, 132)])
def test_forward_matches_torch_reference(dnmod, channels, h, w):
    # partial convolution tiles at every level: 12 x 20 pools to 3 x 5,
    # 68 x 132 leaves a partial 64-pixel tile and an odd row pair
    rng = np.random.default_rng(0)
    dn = dnmod.Denoiser(w, h, channels=channels, seed=3)
    p = dn.get_params()
    # a non-trivial output layer (the default starts near zero)
    p = p + rng.normal(0, 0.02, p.size).astype(np.float32)
    dn.set_params(p)
    x = _rand_input(rng, h, w, channels)
    got, ms = dn.infer(x)
    with torch.no_grad():
        want, _ = _torch_forward(p, x, channels)
    want = want.numpy()
    err = np.abs(got - want)
    scale = np.abs(want - x[..., :3]).max()
    print(f"[denoise] forward: max |err| {err.max():.3e}, change scale {scale:.3e}, {ms:.3f} ms", flush=True)
    assert scale > 1e-2
    # bf16 storage of activations: a stored value can round the other way
    # after a different fp32 summation order (1 bf16 ulp, 2^-8 relative)
    assert err.max() < 2e-2 * scale
    assert np.mean(err) < 2e-3 * scale
    dn.close()


@pytest.mark.parametrize("channels,h,w", [(4, 12, 20), (4, 68, 132), (10, 36, 64)])
def test_fused_inference_equals_unfused(dnmod, channels, h, w):
    """The fused inference kernels (pooling in the convolution epilogue,
    upsample-and-concatenate read in place, the filter applied from the
    logits in LDS) store exactly what the separate kernels store, call after
    call (new inputs, new weights, unfused runs in between)."""
    rng = np.random.default_rng(4)
    dn = dnmod.Denoiser(w, h, channels=channels, seed=9)
    dn.set_params(dn.get_params() + rng.normal(0, 0.02, dn.num_params()).astype(np.float32))
    for rep in range(3):
        x = _rand_input(rng, h, w, channels)
        if rep == 2:
            dn.set_params(dn.get_params() * np.float32(0.9))
        fused, _ = dn.infer(x)
        dn.set_fused(False)
        plain, _ = dn.infer(x)
        dn.set_fused(True)
        print(f"[denoise] fused vs unfused {h}x{w}x{channels} call {rep}: "
              f"max |d| {np.abs(fused - plain).max():.3e}", flush=True)
        np.testing.assert_array_equal(fused.view(np.uint32), plain.view(np.uint32))
    dn.close()


@pytest.mark.parametrize("h,w", [(16, 24), (20, 28)])
def test_gradient_matches_torch_reference(dnmod, h, w):
    rng = np.random.default_rng(1)
    dn = dnmod.Denoiser(w, h, seed=5)
    p = dn.get_params() + rng.normal(0, 0.02, dn.num_params()).astype(np.float32)
    dn.set_params(p)
    x = _rand_input(rng, h, w)
    t = np.clip(x[..., :3] + rng.normal(0, 0.05, (h, w, 3)), 0, None).astype(np.float32)
    g, loss = dn.gradient(x, t)
    pred, layers = _torch_forward(p, x)
    tl = torch.mean((pred - torch.tensor(t)) ** 2)
    tl.backward()
    assert abs(loss - tl.item()) < 1e-2 * tl.item()
    o = 0
    for i, ((cin, cout, _), (wt, bt, _)) in enumerate(zip(_layers(4), layers)):
        gw = g[o:o + 9 * cout * cin].reshape(3, 3, cout, cin).transpose(2, 3, 0, 1)
        o += 9 * cout * cin
        gb = g[o:o + cout]
        o += cout
        rw = wt.grad.numpy()
        rb = bt.grad.numpy()
        cos = float(np.dot(gw.ravel(), rw.ravel()) / (np.linalg.norm(gw) * np.linalg.norm(rw) + 1e-30))
        rel = float(np.linalg.norm(gw - rw) / (np.linalg.norm(rw) + 1e-30))
        relb = float(np.linalg.norm(gb - rb) / (np.linalg.norm(rb) + 1e-30))
        print(f"[denoise] layer {i}: cos {cos:.5f} rel {rel:.3e} bias rel {relb:.3e}", flush=True)
        # masked gradients are rounded to bf16 before the matrix products
        assert cos > 0.999 and rel < 5e-2 and relb < 5e-2, i
    dn.close()


def test_training_reduces_loss(dnmod):
    rng = np.random.default_rng(2)
    h = w = 32
    n = 8
    yy, xx = np.mgrid[0:h, 0:w] / 32.0
    ts, xs = [], []
    for i in range(n):
        f = rng.uniform(1, 4)
        clean = np.stack([0.5 + 0.4 * np.sin(f * xx + i), 0.5 + 0.4 * np.cos(f * yy),
                          0.5 + 0.3 * np.sin(f * (xx + yy))], -1).astype(np.float32)
        x = _rand_input(rng, h, w)
        x[..., :3] = clean + rng.normal(0, 0.15, clean.shape)
        xs.append(x)
        ts.append(clean)
    xs, ts = np.stack(xs), np.stack(ts)
    dn = dnmod.Denoiser(w, h, seed=9)
    losses = dnmod.train(dn, xs, ts, steps=120, batch=4, lr=2e-3)
    print(f"[denoise] loss {np.mean(losses[:5]):.4e} -> {np.mean(losses[-10:]):.4e}", flush=True)
    # the untrained network already filters with a Gaussian; training must
    # still find a clearly better filter for these smooth images
    assert np.mean(losses[-10:]) < 0.7 * np.mean(losses[:5])
    dn.close()


def test_aov_buffers():
    from mdwtracker_amd import capi
    sc = scenes.make_scene("ndda_cpu_2ch64", seed=4)
    a = capi.MdwtRenderer(0)
    b = capi.MdwtRenderer(0)
    for r in (a, b):
        scenes.load_scene(r, sc, 2)
        r.set_frame(64, 48, 2, True, (0.0, 0.0, 0.0))
    b.set_aov(True)
    a.render(3)
    b.render(3)
    np.testing.assert_array_equal(a.read_accum(), b.read_accum())
    aov = b.read_aov()
    assert aov.shape == (48, 64, 4)  # depth, transmittance, two channels
    s = aov[..., 1] + aov[..., 2:].sum(-1)
    np.testing.assert_allclose(s, 1.0, atol=1e-5)
    hit = 1.0 - aov[..., 1]
    assert hit.max() > 0.3 and np.all(aov[..., 0][hit == 0] == 0)
    # a frame limit: the buffers hold the means over the first 2 frames (the
    # same frame ids as a 2-frame run) and the image is unchanged
    c, d = capi.MdwtRenderer(0), capi.MdwtRenderer(0)
    for x in (c, d):
        scenes.load_scene(x, sc, 2)
        x.set_frame(64, 48, 2, True, (0.0, 0.0, 0.0))
        x.set_aov(True)
    c.set_aov_frames(2)
    c.render(3)
    d.render(2)
    np.testing.assert_array_equal(c.read_accum(), a.read_accum())
    np.testing.assert_array_equal(c.read_aov(), d.read_aov())
    with pytest.raises(capi.MdwtError):
        c.set_aov_frames(-1)
    for x in (a, b, c, d):
        x.close()


def test_aov_buffers_ten_channels():
    from mdwtracker_amd import capi
    sc = scenes.make_scene("large_10ch_f32", seed=6, scale=8)
    r = capi.MdwtRenderer(0)
    scenes.load_scene(r, sc, 2)
    r.set_frame(64, 64, 1, True, (0.0, 0.0, 0.0))
    r.set_aov(True)
    r.render(8)
    aov = r.read_aov()
    assert aov.shape == (64, 64, 12)
    np.testing.assert_allclose(aov[..., 1] + aov[..., 2:].sum(-1), 1.0, atol=1e-5)
    frac = aov[..., 2:].sum((0, 1))
    print(f"[denoise] ten-channel hit fractions {np.round(frac / frac.sum(), 3)}", flush=True)
    assert np.count_nonzero(frac[4:]) >= 3  # channels past the fourth are recorded
    r.close()


def test_train_step_weights(dnmod):
    from mdwtracker_amd import capi
    rng = np.random.default_rng(3)
    h = w = 16
    xs = np.stack([_rand_input(rng, h, w) for _ in range(2)])
    ts = np.clip(xs[..., :3] + rng.normal(0, 0.1, (2, h, w, 3)), 0, None).astype(np.float32)
    a = dnmod.Denoiser(w, h, seed=11)
    b = dnmod.Denoiser(w, h, seed=11)
    c = dnmod.Denoiser(w, h, seed=11)
    p0 = a.get_params()
    # the gradient is that of sum_i w_i MSE_i / count: weights (2, 0) on two
    # pairs give exactly the gradient of the first pair alone (powers of two)
    la = a.train_step(xs, ts, 1e-3, np.array([2.0, 0.0], np.float32))
    lb = b.train_step(xs[:1], ts[:1], 1e-3)
    pa, pb = a.get_params(), b.get_params()
    assert np.abs(pa - p0).max() > 1e-4
    # gradients are reduced in a fixed order (no atomics), the zero-weight
    # pair adds exact zeros: the steps are identical bit for bit
    np.testing.assert_array_equal(pa, pb)
    assert lb > 0 and la > 0  # the reported loss is the unweighted mean
    # all weights zero: no step at all
    c.train_step(xs, ts, 1e-3, np.zeros(2, np.float32))
    np.testing.assert_array_equal(c.get_params(), p0)
    with pytest.raises(capi.MdwtError):
        c.train_step(xs, ts, 1e-3, np.array([1.0, -1.0], np.float32))
    for d in (a, b, c):
        d.close()


def test_training_is_deterministic(dnmod):
    rng = np.random.default_rng(7)
    h = w = 32
    xs = np.stack([_rand_input(rng, h, w) for _ in range(4)])
    ts = np.clip(xs[..., :3] + rng.normal(0, 0.1, (4, h, w, 3)), 0, None).astype(np.float32)
    out = []
    for _ in range(2):
        d = dnmod.Denoiser(w, h, seed=13)
        for i in range(3):
            d.train_step(xs[i:i + 2], ts[i:i + 2], 1e-3, np.array([1.0, 0.5], np.float32))
        out.append(d.get_params())
        d.close()
    np.testing.assert_array_equal(out[0], out[1])


def test_accumulated_inputs_and_error_estimate(dnmod):
    from mdwtracker_amd import capi
    sc = scenes.make_scene("sync_4ch256", seed=21, scale=4)
    r, q = capi.MdwtRenderer(0), capi.MdwtRenderer(0)
    for x in (r, q):
        scenes.load_scene(x, sc, 2)
        x.set_frame(64, 64, 1, True, (0.0, 0.0, 0.0))
    levels = (1, 4, 16, 64)
    xs = list(dnmod.accumulate_inputs(r, levels, 9))
    # the last read equals one uninterrupted run of the same seed and frame
    # ids (the frame id, like the reference's frameID, is not rewound by a
    # reset of the accumulation, so both renderers start from a fresh init)
    q.set_seed(9)
    q.reset_accumulation()
    q.render(levels[-1])
    np.testing.assert_array_equal(xs[-1][..., :3], q.read_accum()[..., :3])
    q.close()
    ref, ref_var = dnmod.render_mean(r, 2048, 5)
    sigma2 = float(ref_var.mean() * 2048)
    assert sigma2 > 1e-3
    prev_se = None
    for k, x in zip(levels, xs):
        assert x.shape == (64, 64, dnmod.INPUTS) and np.all(np.isfinite(x))
        assert np.all(x[..., 10] == np.float32(np.log2(k) / 16))
        est, se = dnmod.noise_corrected_mse(x[..., :3], ref, ref_var)
        model = sigma2 / k
        print(f"[denoise] k={k}: plain mse {est:.3e} +- {se:.1e}, model {model:.3e}", flush=True)
        assert abs(est - model) < 0.25 * model + 4 * se
        if k > 1:
            mse_se = float(np.mean(x[..., 9]))
            if prev_se is not None:
                assert mse_se < prev_se
            prev_se = mse_se
    r.close()
