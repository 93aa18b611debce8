"""The surrogate's MLP: FP64 oracle pinned by finite differences (CPU), device forward, gradient
and training against the oracle (GPU).

Device bars: forward and gradient within 1e-3 relative (norm) of the oracle run with the same
bfloat16 rounding points (the only remaining difference is FP32 accumulation order inside the
MFMA tiles); Adam training on the device reduces the loss of a smooth regression by > 10x."""
import numpy as np
import pytest

from oracle import mlp_ref as R

D_IN, H, LAYERS, D_OUT = 30, 128, 3, 3


def random_params(seed, d_in=D_IN, h=H, layers=LAYERS, d_out=D_OUT):
    rng = np.random.default_rng(seed)
    dims = [d_in] + [h] * layers + [d_out]
    return [(rng.normal(0, 1.0 / np.sqrt(i), (i, o)), rng.normal(0, 0.1, o))
            for i, o in zip(dims[:-1], dims[1:])]


def test_oracle_gradient_matches_finite_differences():
    params = random_params(1, d_in=5, h=8, layers=2, d_out=3)
    rng = np.random.default_rng(2)
    x = rng.normal(size=(7, 5))
    t = rng.normal(size=(7, 3))
    loss, g = R.loss_and_grad(params, x, t)
    p0 = R.pack(params)
    h = 1e-6
    for k in rng.choice(p0.size, 40, replace=False):
        pp, pm = p0.copy(), p0.copy()
        pp[k] += h
        pm[k] -= h
        lp, _ = R.loss_and_grad(R.unpack(pp, 5, 8, 2, 3), x, t)
        lm, _ = R.loss_and_grad(R.unpack(pm, 5, 8, 2, 3), x, t)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 + 1e-6 * abs(fd), k


def test_gelu_derivative_and_bf16_rounding():
    x = np.linspace(-6, 6, 1001)
    fd = (R.gelu(x + 1e-6) - R.gelu(x - 1e-6)) / 2e-6
    np.testing.assert_allclose(R.gelu_grad(x), fd, atol=1e-8)
    # bfloat16 keeps 8 significant bits, rounds ties to even
    assert R.bf16(1.0 + 2 ** -8) == 1.0
    assert R.bf16(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6
    assert R.bf16(np.float32(3.0)) == 3.0


@pytest.mark.gpu
def test_device_forward_matches_oracle():
    from nbodygo_amd.mlp import MLP
    params = random_params(3)
    rng = np.random.default_rng(4)
    x = rng.normal(size=(1000, D_IN)).astype(np.float32)
    with MLP(D_IN, H, LAYERS, D_OUT) as m:
        m.set_params(R.pack(params).astype(np.float32))
        y = m.forward(x)
    p32 = R.unpack(R.pack(params).astype(np.float32), D_IN, H, LAYERS, D_OUT)
    exp, _ = R.forward(p32, x, bf=True)
    assert np.linalg.norm(y - exp) / np.linalg.norm(exp) < 1e-3
    nobf, _ = R.forward(p32, x, bf=False)
    assert np.linalg.norm(y - nobf) / np.linalg.norm(nobf) < 3e-2  # bf16 itself


@pytest.mark.gpu
@pytest.mark.parametrize("d_in,hidden,layers,rows", [(17, 128, 1, 1000), (30, 128, 3, 4099),
                                                     (30, 64, 2, 77), (30, 128, 3, 70001),
                                                     (30, 64, 4, 130)])
def test_fused_inference_equals_layer_by_layer_forward(d_in, hidden, layers, rows):
    """The surrogate's one-kernel inference reproduces the GEMM-by-GEMM forward bit for bit
    (same MFMA sequence, same epilogue), including ragged row counts.  Up to three hidden
    layers run the persistent kernel (70,001 rows: several tiles per workgroup); four run the
    one-tile-per-workgroup kernel."""
    from nbodygo_amd.mlp import MLP
    rng = np.random.default_rng(rows)
    x = rng.normal(size=(rows, d_in)).astype(np.float32)
    with MLP(d_in, hidden, layers, D_OUT, seed=rows) as m:
        np.testing.assert_array_equal(m.infer(x), m.forward(x))
        x2 = rng.normal(size=(256, d_in)).astype(np.float32)
        t2 = rng.normal(size=(256, D_OUT)).astype(np.float32)
        m.train(x2, t2, steps=3, batch=256)  # weights move: the transposed copies follow
        np.testing.assert_array_equal(m.infer(x), m.forward(x))


@pytest.mark.gpu
def test_device_gradient_matches_oracle():
    from nbodygo_amd.mlp import MLP
    params = random_params(5)
    rng = np.random.default_rng(6)
    x = rng.normal(size=(512, D_IN)).astype(np.float32)
    t = rng.normal(size=(512, D_OUT)).astype(np.float32)
    p = R.pack(params).astype(np.float32)
    with MLP(D_IN, H, LAYERS, D_OUT) as m:
        m.set_params(p)
        np.testing.assert_array_equal(m.get_params(), p)
        g, loss = m.grad(x, t)
    l_ref, g_ref = R.loss_and_grad(R.unpack(p, D_IN, H, LAYERS, D_OUT), x, t, bf=True)
    assert loss == pytest.approx(l_ref, rel=1e-3)
    assert np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref) < 1e-3
    # each layer's block, not only the total
    off = 0
    for W, b in R.unpack(p, D_IN, H, LAYERS, D_OUT):
        k = W.size + b.size
        seg, ref = g[off:off + k], g_ref[off:off + k]
        assert np.linalg.norm(seg - ref) <= 2e-3 * np.linalg.norm(ref) + 1e-12
        off += k


@pytest.mark.gpu
def test_device_training_reduces_loss():
    from nbodygo_amd.mlp import MLP
    rng = np.random.default_rng(7)
    x = rng.uniform(-1, 1, size=(65536, D_IN)).astype(np.float32)
    w = rng.normal(size=(D_IN, D_OUT))
    t = np.tanh(x @ w / 3.0).astype(np.float32) + 0.1 * np.sin(3 * x[:, :3])
    with MLP(D_IN, H, LAYERS, D_OUT, seed=1) as m:
        y0 = m.forward(x[:4096])
        l0 = float(np.mean((y0 - t[:4096]) ** 2))
        lt = m.train(x, t, steps=600, batch=1024, lr=2e-3)
        y1 = m.forward(x[:4096])
        l1 = float(np.mean((y1 - t[:4096]) ** 2))
    assert l1 < l0 / 10 and lt < l0 / 10


# The two tests below run torch on the CPU.  They come after this file's GPU tests: in a
# process whose first GPU use comes after torch's autograd has started, hipGetDeviceCount has
# been seen to report no device on the GPU boxes, so a run of this file alone keeps the GPU
# tests first.
def test_oracle_forward_and_gradient_match_torch():
    """mlp_ref against an independent statement of the same network: torch's tanh-form GELU,
    its linear layers and autograd, in float64 (no bf16 rounding), so a wrong GELU constant, a
    dropped bias or a transposed weight in the oracle fails here."""
    torch = pytest.importorskip("torch")
    F = torch.nn.functional
    params = random_params(5, d_in=6, h=16, layers=3, d_out=3)
    rng = np.random.default_rng(6)
    x = rng.normal(size=(33, 6))
    t = rng.normal(size=(33, 3))
    ws = [(torch.tensor(W, requires_grad=True), torch.tensor(b, requires_grad=True))
          for W, b in params]
    h = torch.tensor(x)
    for W, b in ws[:-1]:
        h = F.gelu(F.linear(h, W.T, b), approximate="tanh")
    W, b = ws[-1]
    y = F.linear(h, W.T, b)
    loss = torch.mean((y - torch.tensor(t)) ** 2)
    loss.backward()
    y_or, _ = R.forward(params, x)
    np.testing.assert_allclose(y_or, y.detach().numpy(), rtol=1e-12, atol=1e-12)
    l_or, g_or = R.loss_and_grad(params, x, t)
    assert abs(l_or - loss.item()) <= 1e-12 * abs(loss.item())
    g_t = R.pack([(W.grad.numpy(), b.grad.numpy()) for W, b in ws])
    np.testing.assert_allclose(g_or, g_t, rtol=1e-10, atol=1e-13)
    z = np.linspace(-8, 8, 2001)
    np.testing.assert_allclose(R.gelu(z), F.gelu(torch.tensor(z), approximate="tanh").numpy(),
                               rtol=1e-13, atol=1e-15)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(7)
    v = np.concatenate([rng.normal(size=20000) * 10.0 ** rng.integers(-20, 20, 20000),
                        np.float32(1.0) + np.arange(-512, 512) * np.float32(2.0 ** -16)])
    v32 = v.astype(np.float32)
    ref = torch.tensor(v32).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(R.bf16(v32), ref)
