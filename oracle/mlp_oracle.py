"""TEST INFRASTRUCTURE ONLY -- CPU restatement of the transmission surrogate (no reference
counterpart): an MLP n_in -> width (x depth, GELU) -> n_out with the GPU's quantisation points.

With bf16=True every GEMM operand (weights, hidden activations, back-propagated errors) is rounded
to bfloat16 (round to nearest even, as the device does) and the products are summed in FP64;
hidden pre-activations, outputs and gradients stay unrounded.  With bf16=False the same code is the
plain FP64 network, whose gradient is pinned by central finite differences in tests/.

Parameter layout (FP32 vector): per layer l, W_l [fout][fin] row-major then b_l [fout].
Loss: mean over samples and outputs of (y - t)^2.
"""
import numpy as np
from scipy.special import erf


def bf16(x):
    """Round float32 values to bfloat16 (nearest even) and return them as float64."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def layout(n_in, width, depth, n_out):
    """[(fin, fout, w_offset, b_offset)] per layer and the parameter count."""
    out, off = [], 0
    for l in range(depth + 1):
        fin = n_in if l == 0 else width
        fout = n_out if l == depth else width
        out.append((fin, fout, off, off + fin * fout))
        off += fin * fout + fout
    return out, off


def init_params(n_in, width, depth, n_out, seed):
    """He-uniform weights, zero biases (FP32)."""
    rng = np.random.default_rng(seed)
    lay, n = layout(n_in, width, depth, n_out)
    p = np.zeros(n, np.float32)
    for fin, fout, wo, bo in lay:
        lim = np.sqrt(6.0 / fin)
        p[wo:wo + fin * fout] = rng.uniform(-lim, lim, fin * fout).astype(np.float32)
    return p


def _q(x, on):
    return bf16(x) if on else np.asarray(x, np.float64)


def forward(p, x, dims, bf16_ops=True):
    """Returns (y, cache) for inputs x [n][n_in]."""
    lay, _ = layout(*dims)
    p = np.asarray(p, np.float32) if bf16_ops else np.asarray(p, np.float64)
    a = _q(x, bf16_ops)
    cache = []
    for l, (fin, fout, wo, bo) in enumerate(lay):
        W = _q(p[wo:wo + fin * fout].reshape(fout, fin), bf16_ops)
        z = a @ W.T + p[bo:bo + fout].astype(np.float64)
        cache.append((a, W, z))
        if l < len(lay) - 1:
            a = _q(gelu(z).astype(np.float32) if bf16_ops else gelu(z), bf16_ops)
    return z, cache


def loss_grad(p, x, t, dims, bf16_ops=True):
    """Loss and gradient (parameter layout, FP64) with the device's quantisation points."""
    lay, n = layout(*dims)
    y, cache = forward(p, x, dims, bf16_ops)
    t = np.asarray(t, np.float64)
    loss = float(np.mean((y - t) ** 2))
    g = np.zeros(n)
    dz = 2.0 * (y - t) / y.size
    for l in range(len(lay) - 1, -1, -1):
        fin, fout, wo, bo = lay[l]
        a, W, z = cache[l]
        dzq = _q(dz.astype(np.float32) if bf16_ops else dz, bf16_ops)
        g[wo:wo + fin * fout] = (dzq.T @ a).reshape(-1)
        g[bo:bo + fout] = dz.sum(axis=0)
        if l == 0:
            break
        dx = dzq @ W
        dz = dx * gelu_grad(cache[l - 1][2])
    return loss, g


def adam_update(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8, bias_correction=True):
    """One Adam step (Kingma & Ba 2015, Algorithm 1) in FP32, in the order the device evaluates it:
    m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2, p -= lr (m / c1) / (sqrt(v / c2) + eps) with
    c1 = 1 - b1^step, c2 = 1 - b2^step.  Returns (p, m, v).  bias_correction=False drops c1 and c2
    (a planted mistake for the tests)."""
    f = np.float32
    p, g, m, v = (np.asarray(a, np.float32) for a in (p, g, m, v))
    m = f(b1) * m + (f(1) - f(b1)) * g
    v = f(b2) * v + (f(1) - f(b2)) * g * g
    c1 = f(1) - f(np.power(f(b1), f(step))) if bias_correction else f(1)
    c2 = f(1) - f(np.power(f(b2), f(step))) if bias_correction else f(1)
    p = p - f(lr) * (m / c1) / (np.sqrt(v / c2) + f(eps))
    return p.astype(np.float32), m, v
