"""CPU check of the taper family's device-raster inputs (3d_fdtd_amd/design.py): the material
table plus per-(k, i) ids that fdtd_dev_set_materials_extruded expands on the device describe
exactly the float32 permittivity volume of the host raster, for the design-loop family and the
620 x 200 x 100 taper (where the 500 fill-fraction planes need 16-bit ids)."""
import importlib

import numpy as np
import pytest

design = importlib.import_module("3d_fdtd_amd.design")


def families():
    return {"design-loop": design.TaperFamily(),
            "taper-620": design.TaperFamily(disc=0.02, I=200, J=100, core_cells=11, npml=10, port=48, steps=10,
                                            lam=(1.55, 1.55), nfreq=1)}


@pytest.mark.parametrize("name", ["design-loop", "taper-620"])
def test_extruded_raster_equals_host_raster(name):
    fam = families()[name]
    rng = np.random.default_rng(41)
    designs = list(fam.random(rng, 12)) + [np.concatenate([np.linspace(0.5, 3.0, 8), [10.0]])]
    for p in designs:
        ep, K = fam.permittivity(p)
        (table, ids, j0, j1, bg), K2 = fam.extruded(p)
        assert K2 == K and ids.shape == (K, fam.I) and table.shape[1] == 3
        assert ids.max() < table.shape[0] <= 65536
        vol = np.empty((K, fam.J, fam.I), np.float32)
        vol[:] = table[bg, 0]
        vol[:, j0:j1, :] = table[ids, 0][:, None, :]
        host = ep.reshape(K, fam.J, fam.I, 3)
        for comp in range(3):
            assert np.array_equal(vol, host[..., comp])
        # the table holds each value once, and the cladding is the background
        assert len(np.unique(table[:, 0])) == table.shape[0]
        assert table[bg, 0] == np.float32(fam.eps_clad)
