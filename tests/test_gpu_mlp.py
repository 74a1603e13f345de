"""GPU tests of the transmission surrogate (BF16 MFMA MLP kernels, include/fdtd3d_mlp.h) against
its CPU oracle (oracle/mlp_oracle.py, same BF16 quantisation points, FP64 sums; its backward pass
is pinned by finite differences in test_oracle_mlp.py).

Gates: forward outputs relative L2 <= 2e-3 and every layer's gradient relative L2 <= 1e-2 (FP32
matrix-core sums against FP64 ones, and the occasional BF16 rounding flip that follows from them);
a 10^5-sample inference agrees with the small-batch one row for row; training on a smooth
synthetic map cuts the loss by at least 10x.
"""
import importlib
import sys

import numpy as np
import pytest

from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu

fd = importlib.import_module("3d_fdtd_amd")
DIMS = (9, 256, 4, 32)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def _case(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, DIMS[0])).astype(np.float32)
    t = rng.uniform(-1, 1, (n, DIMS[3])).astype(np.float32)
    return x, t


def test_forward_and_gradients_match_oracle():
    p = M.init_params(*DIMS, seed=1)
    x, t = _case(37, 2)
    net = fd.Mlp(*DIMS, params=p)
    y = net.forward(x)
    y_o, _ = M.forward(p, x, DIMS)
    ef = rel(y, y_o)
    loss, g = net.loss_grad(x, t)
    loss_o, g_o = M.loss_grad(p, x, t, DIMS)
    say("forward rel L2 %.2e, loss %.6e vs %.6e" % (ef, loss, loss_o))
    assert ef <= 2e-3
    assert abs(loss - loss_o) <= 1e-3 * abs(loss_o)
    lay, _ = M.layout(*DIMS)
    for l, (fin, fout, wo, bo) in enumerate(lay):
        ew = rel(g[wo:wo + fin * fout], g_o[wo:wo + fin * fout])
        eb = rel(g[bo:bo + fout], g_o[bo:bo + fout])
        say("layer", l, "dW rel L2 %.2e, db rel L2 %.2e" % (ew, eb))
        assert ew <= 1e-2 and eb <= 1e-2
    net.free()


def test_large_batch_inference_consistent():
    p = M.init_params(*DIMS, seed=4)
    net = fd.Mlp(*DIMS, params=p)
    x, _ = _case(100000, 5)
    y = net.forward(x)
    y_small = net.forward(x[:37])
    assert np.array_equal(y[:37], y_small)
    assert np.isfinite(y).all()
    net.free()


def test_training_reduces_loss():
    rng = np.random.default_rng(9)
    x = rng.uniform(-1, 1, (1024, DIMS[0])).astype(np.float32)
    A = rng.normal(0, 0.6, (DIMS[0], DIMS[3]))
    t = np.sin(x @ A).astype(np.float32)  # a smooth map of the 9 inputs
    net = fd.Mlp(*DIMS, params=M.init_params(*DIMS, seed=3))
    first = net.adam_step(x, t, lr=1e-3)
    for _ in range(300):
        last = net.adam_step(x, t, lr=1e-3)
    say("training loss %.4e -> %.4e" % (first, last))
    assert last < 0.1 * first
    net.free()


def test_resident_training_equals_step_by_step():
    """fdtd_mlp_train (data uploaded once, loss read back only at the ends) performs exactly the
    arithmetic of the same number of fdtd_mlp_adam_step calls: parameters equal bit for bit."""
    x, t = _case(300, 12)
    a = fd.Mlp(*DIMS, params=M.init_params(*DIMS, seed=8))
    b = fd.Mlp(*DIMS, params=M.init_params(*DIMS, seed=8))
    first_a = a.adam_step(x, t, lr=2e-3)
    for _ in range(14):
        a.adam_step(x, t, lr=2e-3)
    first_b, last_b = b.train(x, t, 15, lr=2e-3)
    pa, pb = a.params(), b.params()
    say("losses: first %.6e / %.6e, after %.6e; max |dp| %.1e" % (first_a, first_b, last_b, np.abs(pa - pb).max()))
    assert first_a == first_b
    assert np.array_equal(pa, pb)
    assert last_b < first_b
    a.free()
    b.free()


def test_gradients_match_oracle_at_training_batch():
    """The training batch of 4096 designs: the weight-gradient GEMMs split K = 4096 over many
    blocks and reduce the partial sums in a fixed order; every layer's gradient still equals the
    oracle's FP64 sums to the gate."""
    p = M.init_params(*DIMS, seed=11)
    x, t = _case(4096, 12)
    net = fd.Mlp(*DIMS, params=p)
    loss, g = net.loss_grad(x, t)
    loss_o, g_o = M.loss_grad(p, x, t, DIMS)
    say("batch 4096: loss %.6e vs %.6e" % (loss, loss_o))
    assert abs(loss - loss_o) <= 1e-3 * abs(loss_o)
    lay, _ = M.layout(*DIMS)
    for l, (fin, fout, wo, bo) in enumerate(lay):
        ew = rel(g[wo:wo + fin * fout], g_o[wo:wo + fin * fout])
        eb = rel(g[bo:bo + fout], g_o[bo:bo + fout])
        say("layer", l, "dW rel L2 %.2e, db rel L2 %.2e" % (ew, eb))
        assert ew <= 1e-2 and eb <= 1e-2
    g_again = net.loss_grad(x, t)[1]
    assert np.array_equal(g, g_again)  # deterministic reductions
    net.free()


def test_adam_steps_match_oracle_restatement():
    """Two Adam steps on the GPU (batch 4096) against the FP32 restatement of Adam fed with the GPU's
    own gradients (read with loss_grad right before each step): the parameters agree to the last
    bit or two of FP32.  The first step is also the closed form p - lr g / (|g| + eps)."""
    p0 = M.init_params(*DIMS, seed=13)
    x, t = _case(4096, 14)
    net = fd.Mlp(*DIMS, params=p0)
    lr = 2e-3
    g1 = net.loss_grad(x, t)[1]
    net.adam_step(x, t, lr=lr)
    p1 = net.params()
    g2 = net.loss_grad(x, t)[1]
    net.adam_step(x, t, lr=lr)
    p2 = net.params()
    z = np.zeros(p0.size, np.float32)
    r1, m1, v1 = M.adam_update(p0, g1, z, z, 1, lr)
    r2, _, _ = M.adam_update(r1, g2, m1, v1, 2, lr)
    e1 = float(np.abs(p1.astype(np.float64) - r1).max())
    e2 = float(np.abs(p2.astype(np.float64) - r2).max())
    closed = float(np.abs(p1.astype(np.float64) - (p0 - lr * g1.astype(np.float64) / (np.abs(g1) + 1e-8))).max())
    say("adam: step 1 max |dp| %.2e (closed form %.2e), step 2 %.2e" % (e1, closed, e2))
    assert e1 <= 4e-7 and e2 <= 4e-7 and closed <= 4e-7
    net.free()
