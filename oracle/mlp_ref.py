"""FP64 reference of the device MLP (d_in -> hidden x layers, GELU -> d_out) -- TEST INFRASTRUCTURE.

Forward and backward of the mean-squared-error loss in float64.  With bf16=True the values are
rounded to bfloat16 (round to nearest even) exactly where the device kernels round them: the GEMM
inputs (x, weights, hidden activations, output and hidden gradients); accumulations, biases,
pre-activations and the GELU derivative stay in high precision.  The analytic gradient is pinned
against central finite differences in tests/test_mlp.py.
"""
import numpy as np

K0 = 0.7978845608028654  # sqrt(2/pi)
K1 = 0.044715


def bf16(a):
    """Round float values to bfloat16 (round to nearest even), returned as float64."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(K0 * (x + K1 * x ** 3)))


def gelu_grad(x):
    u = K0 * (x + K1 * x ** 3)
    t = np.tanh(u)
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * K0 * (1.0 + 3.0 * K1 * x * x)


def unpack(p, d_in, hidden, layers, d_out):
    dims = [d_in] + [hidden] * layers + [d_out]
    out, off = [], 0
    for i, o in zip(dims[:-1], dims[1:]):
        W = np.asarray(p[off:off + i * o], np.float64).reshape(i, o)
        off += i * o
        b = np.asarray(p[off:off + o], np.float64)
        off += o
        out.append((W, b))
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers])


def forward(params, x, bf=False):
    """Returns (y, cache)."""
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    h = r(x)
    cache = [h]
    zs = []
    for W, b in params[:-1]:
        z = h @ r(W) + b
        zs.append(z)
        h = r(gelu(z))
        cache.append(h)
    W, b = params[-1]
    y = h @ r(W) + b
    return y, (cache, zs)


def loss_and_grad(params, x, t, bf=False):
    """MSE over all outputs; gradient in the pack() layout."""
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    y, (cache, zs) = forward(params, x, bf)
    n = y.size
    loss = np.sum((y - t) ** 2) / n
    g = 2.0 * (y - t) / n
    grads = [None] * len(params)
    gb = r(g)
    W, b = params[-1]
    grads[-1] = (cache[-1].T @ gb, g.sum(0))
    for l in range(len(params) - 2, -1, -1):
        dh = gb @ r(params[l + 1][0]).T
        dz = dh * gelu_grad(zs[l])
        gb = r(dz)
        grads[l] = (cache[l].T @ gb, dz.sum(0))
    return loss, pack(grads)
