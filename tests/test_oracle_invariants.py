"""Invariants that pin the oracle's extensions (no reference counterpart) -- CPU only.

* Zero-thickness CPML reduces to the reference interior update bit for bit.
* A CPML-terminated dipole run matches a much larger domain (whose walls are too far for any
  reflection to return) in the central region: reflection from the PML is small.  A planted
  sign error in the psi recursion must break this, so the test cannot pass vacuously.
* The multi-frequency scalar DFT of a pure tone has the closed form sum_n e^{i(w-w0)n dt}/2i...
  checked against direct numpy summation of the stated formula on a tone.
"""
import ctypes as C

import numpy as np

from oracle import oracle_py as O


def test_zero_width_cpml_is_reference_interior():
    rng = np.random.default_rng(11)
    I, J, K, disc = 14, 12, 16, 0.02
    dt = 0.99 * disc / np.sqrt(3)
    a, b = O.Grid(I, J, K, disc), O.Grid(I, J, K, disc)
    for g in (a, b):
        g.ep[:] = rng.uniform(1, 12, g.ep.size).astype(np.float32) if g is a else a.ep
        g.sig[:] = rng.uniform(0, 3, g.sig.size).astype(np.float32) if g is a else a.sig
        g.src[:] = rng.uniform(-1, 1, g.src.size).astype(np.float32) if g is a else a.src
        g.f[:] = rng.uniform(-1, 1, g.f.size) if g is a else a.f
    cp = O.Cpml(a, (0, 0, 0, 0, 0, 0), dt)
    for _ in range(10):
        O.step_cpml(a, cp, dt)
        O.lib().orc_update_E(b.ptr, dt)
        O.lib().orc_update_H(b.ptr, dt)
    assert np.array_equal(a.f, b.f)


def _dipole(N, npml, steps, dt, disc, corrupt=False):
    g = O.Grid(N, N, N, disc)
    cp = O.Cpml(g, (npml,) * 6, dt, order=3, R0=1e-8, kappa_max=1.0, alpha_max=0.05)
    if corrupt:
        for a in range(3):
            cp.tables[a][1][:] *= -1.0  # flip the sign of the E psi recursion coefficient
    c0 = N // 2 + N * (N // 2 + N * (N // 2))
    w0 = 2 * np.pi / 1.55
    out = []
    for n in range(steps):
        t = n * dt
        g.src[6 * c0 + 2] = np.float32(np.exp(-((t - 60 * dt) / (20 * dt)) ** 2) * np.sin(w0 * t))
        O.step_cpml(g, cp, dt)
        out.append(g.comp(2)[N // 2 - 6:N // 2 + 6, N // 2 - 6:N // 2 + 6, N // 2 - 6:N // 2 + 6].copy())
    return np.array(out)


def test_cpml_reflection_small_vs_large_domain():
    disc = 0.02
    dt = 0.99 * disc / np.sqrt(3)
    steps = 140
    small = _dipole(40, 8, steps, dt, disc)
    # large domain: walls 60 cells from the centre, reflections need > 2*60/0.57 = 210 steps to return
    big = _dipole(120, 8, steps, dt, disc)
    err = np.linalg.norm(small - big) / np.linalg.norm(big)
    assert err < 1e-4, err
    # the same small box without PML (bare PEC walls) reflects ~1000x more; a planted sign error
    # in the psi recursion is just as visible
    pec = _dipole(40, 0, steps, dt, disc)
    assert np.linalg.norm(pec - big) / np.linalg.norm(big) > 1e-3
    bad = _dipole(40, 8, steps, dt, disc, corrupt=True)
    assert np.linalg.norm(bad - big) / np.linalg.norm(big) > 1e-3


def test_scalar_dft_pure_tone():
    w0, dt, n = 3.0, 0.01, 500
    om = np.array([2.5, 3.0, 3.5])
    acc = np.zeros(6)
    for s in range(n):
        O.lib().orc_dft_scalar(3, O.dptr(om), s * dt, np.cos(w0 * s * dt), O.dptr(acc))
    t = np.arange(n) * dt
    want = np.array([np.sum(np.cos(w0 * t) * np.exp(1j * w * t)) for w in om])
    assert np.allclose(acc[0::2] + 1j * acc[1::2], want, rtol=1e-12, atol=1e-9)
    # resonance: the on-tone bin grows like n/2
    assert abs(acc[2] + 1j * acc[3]) > 0.45 * n
