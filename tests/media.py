"""Small seeded test problems shared by the parity tests."""
import numpy as np

import cudamcfl_amd as M

# rows: n, mua, mus, flc, AbsCExi, AbsCEmi, eY, g   (1/cm, .mci medium-line order)
MATS3 = [
    [1.4, 0.5, 50.0, 1.0, 0.1, 0.08, 0.1, 0.9],
    [1.4, 2.0, 80.0, 1.0, 0.5, 0.3, 0.1, 0.8],
    [1.5, 0.3, 30.0, 0.0, 0.0, 0.0, 0.0, 0.0],
]


def small_labels(nx=20, ny=20, nz=10, seed=0):
    """Background 1, a sphere of 2, a lower slab of 3 (index step 1.4 -> 1.5)."""
    rng = np.random.default_rng(seed)
    lab = np.ones((nz, ny, nx), np.uint8)
    z, y, x = np.mgrid[0:nz, 0:ny, 0:nx]
    c = rng.uniform([nz * 0.3, ny * 0.3, nx * 0.3], [nz * 0.6, ny * 0.7, nx * 0.7])
    lab[(z - c[0]) ** 2 + (y - c[1]) ** 2 + (x - c[2]) ** 2 < (nz * 0.25) ** 2] = 2
    lab[int(nz * 0.7):, :, :] = 3
    return lab


def small_sim(photons=20000, temporal=True, fhd=True, **kw):
    lab = small_labels()
    tmp = (0.0, 0.0, 2, 1, 0.4, 0.4, 0.3, 0, 50, 1000000) if temporal else None
    return M.Simulation.voxel(lab, MATS3, esp=1.0, grid_size=10, det=(0.1, 0.1, 20, 20, 0.0, 0.0),
                              photons=photons, source=(0.05, -0.02, 0.0), n_above=1.0, n_below=1.33,
                              fhd=fhd, temporal=tmp, **kw)
