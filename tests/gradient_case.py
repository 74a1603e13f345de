"""A small forward / adjoint problem shared by the DFT-gradient tests (CPU oracle and GPU).

Grid 20 x 20 x 36 (library axes, z = propagation), 20 nm cells, 5-cell CPML on every face,
background eps 2.0 with a dielectric block eps 6.0.  A smooth synthetic complex "mode" (10 x 10
window, every component a different complex Gaussian) feeds a pulsed mode source at z = 7 and an
overlap monitor at z = 28; the design box [6,14) x [6,14) x [12,20) lies between them.  The mode
does not need to be an eigenmode: the adjoint identity checked here holds for any monitor weights.
"""
import numpy as np

I, J, K = 20, 20, 36
DISC = 0.02
DT = 0.99 * DISC / np.sqrt(3)
NPML = (5, 5, 5, 5, 5, 5)
CPML = dict(order=3, R0=1e-8, kappa_max=1.0, alpha_max=0.05)
OMEGA = 2 * np.pi / 1.55
NSTEPS = 2200
SRC_K, MON_K = 7, 28
BOX = (6, 14, 6, 14, 12, 20)  # i0, i1, j0, j1, k0, k1
MI = MJ = 10


def envelope(n):
    """Gaussian envelope (120 steps wide, centred at step 480).  Its spectrum at omega (the DC
    content of the modulated source, which would leave static charge) and at 2 omega (the image
    term of a real source) is below 1e-6."""
    return np.exp(-((n - 480.0) / 120.0) ** 2 / 2)


def materials():
    ep = np.full((K, J, I, 3), 2.0, np.float32)
    ep[10:22, 8:12, 7:13, :] = 6.0
    return ep.reshape(-1)


def mode_array():
    """Interleaved complex [mj][mi][6] mode, each component a smooth complex Gaussian."""
    y, x = np.mgrid[0:MJ, 0:MI].astype(np.float64)
    out = np.zeros((MJ, MI, 6, 2))
    for m in range(6):
        cx, cy, w = 4.5 + 0.3 * m, 4.5 - 0.2 * m, 2.5 + 0.2 * m
        g = np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * w * w))
        ph = 0.4 * m + 0.15 * x - 0.1 * y
        amp = 1.0 + 0.25 * m
        out[..., m, 0] = amp * g * np.cos(ph)
        out[..., m, 1] = amp * g * np.sin(ph)
    return out.reshape(-1)


def box_cells():
    i0, i1, j0, j1, k0, k1 = BOX
    return (i1 - i0) * (j1 - j0) * (k1 - k0)


def block_mask(corner, size=2):
    """Boolean [k][j][i] mask over the design box of a size^3 block at box-relative corner (i, j, k)."""
    i0, i1, j0, j1, k0, k1 = BOX
    m = np.zeros((k1 - k0, j1 - j0, i1 - i0), bool)
    ci, cj, ck = corner
    m[ck:ck + size, cj:cj + size, ci:ci + size] = True
    return m


def gradient_weight(E0):
    """d a / d eps(cell) = 2 u / (dt E0) * sum_m F_m A_m with u = e^{-i omega dt} - 1."""
    u = np.exp(-1j * OMEGA * DT) - 1.0
    return 2.0 * u / (DT * E0)


BLOCKS = [(1, 1, 1), (4, 2, 3), (2, 5, 6)]
