"""GPU test of the surrogate in the taper design loop (3d_fdtd_amd/design.py): what the trained
network gains over the method without it.

768 random spline tapers (reduced family: 50 nm cells, 48 x 20 cross-section, 1.5-4 um long) are
simulated by batched FDTD; the BF16-MFMA surrogate is trained on them on the GPU.  From a pool of
8192 new candidates, the 48 the surrogate ranks highest and 48 picked without it (the first 48 of
the pool) are verified by full FDTD, which decides every design.  With the same FDTD
budget, the surrogate's picks must transmit clearly more (mean band |t|^2 higher by >= 0.1) and
contain at least twice as many designs above the spec |t|^2 >= 0.7.
"""
import importlib
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

design = importlib.import_module("3d_fdtd_amd.design")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_surrogate_screening_beats_unscreened_picks():
    fam = design.TaperFamily()
    rng = np.random.default_rng(2026)
    t0 = time.time()
    train = fam.random(rng, 768)
    t_train = fam.simulate(train)
    t1 = time.time()
    net = design.fit(fam, train, t_train, steps=1500)
    t2 = time.time()
    pool = fam.random(rng, 8192)
    top, pred = design.screen(fam, net, pool, 48)
    t3 = time.time()
    picked = fam.simulate(pool[top])
    plain = fam.simulate(pool[:48])
    t4 = time.time()
    T_sur, T_plain = design.band_transmission(picked), design.band_transmission(plain)
    spec = 0.7
    n_sur, n_plain = int((T_sur >= spec).sum()), int((T_plain >= spec).sum())
    err = np.abs(pred[top] - T_sur)
    say("FDTD training set %.1f s, training %.1f s, screening 8192 %.2f s, verification 2 x 48 %.1f s"
        % (t1 - t0, t2 - t1, t3 - t2, t4 - t3))
    say("mean band |t|^2: surrogate picks %.3f, unscreened picks %.3f; above spec: %d vs %d"
        % (T_sur.mean(), T_plain.mean(), n_sur, n_plain))
    say("surrogate error on its picks: median %.3f, max %.3f" % (np.median(err), err.max()))
    net.free()
    assert T_sur.mean() >= T_plain.mean() + 0.1
    assert n_sur >= max(2 * n_plain, 4)


def test_surrogate_loop_accepts_only_fdtd_verified_designs():
    """Two iterations of the design loop (3d_fdtd_amd/design.py SurrogateLoop) with the fallback
    forced on (min_hit above 1): every accepted design's FDTD band |t|^2 meets the spec and
    re-simulating it gives the same transmission; the fallback ran once per iteration; the
    verified results joined the training data."""
    fam = design.TaperFamily()
    rng = np.random.default_rng(77)
    train = fam.random(rng, 512)
    t_train = fam.simulate(train)
    net = design.fit(fam, train, t_train, steps=600)
    loop = design.SurrogateLoop(fam, net, train, t_train, spec=0.7, pool=20000, verify=32, refit=10,
                                min_hit=1.01, fallback_iters=8, seed=5)
    loop.run(iterations=2)
    say("accepted %d of %d verified, fallbacks %d (accepted %d)"
        % (len(loop.accepted), loop.verified, loop.fallbacks, loop.fallback_accepted))
    assert loop.iterations == 2 and loop.verified == 64 and loop.fallbacks == 2
    assert sum(len(x) for x in loop.x) == 512 + 64
    assert len(loop.accepted) > 0
    T = design.band_transmission(fam.simulate(np.array(loop.accepted)))
    assert (T >= 0.7).all()
    assert np.allclose(T, loop.T_fdtd, rtol=0, atol=1e-6)
    err = loop.surrogate_error()
    say("surrogate max_lambda |dT| on accepted designs: median %.3f" % np.median(err))
    net.free()
