"""CPU pins of the surrogate's oracle (oracle/mlp_oracle.py; no reference counterpart).

* bfloat16 rounding: known answers (round to nearest, ties to even).
* The backward pass in FP64 mode equals central finite differences of the forward pass; a planted
  mistake (GELU' dropped from the chain rule) must fail the same comparison.
"""
import numpy as np

from oracle import mlp_oracle as M


def test_bf16_rounding_known_answers():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, 1.0 + 2.0 ** -7, -2.5, 3.0e-39, 65504.0], np.float32)
    want = np.array([1.0, 1.0, 1.0078125, 1.0078125, -2.5, 0.0, 65536.0])
    got = M.bf16(x)
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0078125 and got[3] == 1.0078125 and got[4] == -2.5
    assert got[6] == 65536.0
    assert abs(got[5]) < 4e-39 and want[5] == 0.0


def test_backward_matches_finite_differences():
    dims = (9, 12, 2, 6)
    rng = np.random.default_rng(3)
    p = M.init_params(*dims, seed=7).astype(np.float64)
    x = rng.uniform(-1, 1, (11, 9))
    t = rng.uniform(-1, 1, (11, 6))
    _, g = M.loss_grad(p, x, t, dims, bf16_ops=False)
    lay, n = M.layout(*dims)

    def loss(q):
        y, _ = M.forward(q, x, dims, bf16_ops=False)
        return float(np.mean((y - t) ** 2))

    idx = rng.choice(n, 40, replace=False)
    h = 1e-4
    fdg = []
    for i in idx:
        qp, qm = p.copy(), p.copy()
        qp[i] += h
        qm[i] -= h
        fdg.append((loss(qp) - loss(qm)) / (2 * h))
    fdg = np.array(fdg)
    err = np.abs(g[idx] - fdg).max() / np.abs(fdg).max()
    assert err < 1e-6, err
    # planted mistake: drop GELU' from the chain rule
    saved = M.gelu_grad
    try:
        M.gelu_grad = lambda z: np.ones_like(z)
        _, gbad = M.loss_grad(p, x, t, dims, bf16_ops=False)
    finally:
        M.gelu_grad = saved
    assert np.abs(gbad[idx] - fdg).max() / np.abs(fdg).max() > 1e-2


def test_adam_restatement_closed_forms():
    """The first Adam step from zero moments moves each parameter by lr g / (|g| + eps), whatever g's
    scale (the bias-corrected moments are g and g^2); with a constant gradient every later step
    does the same.  Dropping the bias correction (a planted mistake) moves the first step by a
    factor (1 - b1) / sqrt(1 - b2) = 3.16 instead."""
    rng = np.random.default_rng(21)
    p0 = rng.uniform(-1, 1, 1000).astype(np.float32)
    g = (rng.normal(size=1000) * 10.0 ** rng.integers(-4, 3, 1000)).astype(np.float32)
    lr = 1e-3
    want = p0.astype(np.float64) - lr * g / (np.abs(g) + 1e-8)
    p1, m, v = M.adam_update(p0, g, np.zeros(1000), np.zeros(1000), 1, lr)
    assert np.abs(p1 - want).max() <= 2e-7
    p, mm, vv = p1, m, v
    for step in range(2, 6):
        p, mm, vv = M.adam_update(p, g, mm, vv, step, lr)
    assert np.abs((p1 - p) - 4 * (p0 - p1)).max() <= 1e-6
    bad, _, _ = M.adam_update(p0, g, np.zeros(1000), np.zeros(1000), 1, lr, bias_correction=False)
    assert np.abs(bad - want).max() > 5e-4
