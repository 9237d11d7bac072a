"""Readers of the output files the C ABI writes in the reference's layouts
(CUDAMCFLio.cu:57-191, CUDAMCFLmain.cu:368-392,604-647): the photon hitting
density (_PHD-Ascii.dat), the R/T and fluorescence images (_Refl.dat,
_Trans.dat, _FlR.dat, _FlT.dat) and the time-of-flight histograms
(_time.dat).  They stand in for the reference's post-processing notebooks
(PHD Splitter.ipynb:48-57, MCImagesProcess.ipynb)."""
import numpy as np


def read_phd(path, reference_axes=False):
    """_PHD-Ascii.dat: a header line "nx<TAB>ny<TAB>nz", then one value per
    voxel, x fastest.  Returns an array of shape (nz, ny, nx).

    reference_axes=True returns the notebook's arrangement instead
    (PHD Splitter.ipynb:55: reshape to (nz, nx, ny), then two axis swaps):
    shape (nx, ny, nz), indexed [iy, ix, iz] when nx == ny."""
    with open(path) as f:
        nx, ny, nz = (int(v) for v in f.readline().split())
        data = np.loadtxt(f, dtype=np.float64, ndmin=1)
    if data.size != nx * ny * nz:
        raise ValueError(f"{path}: {data.size} values for a {nx}x{ny}x{nz} grid")
    if reference_axes:
        return np.swapaxes(np.swapaxes(data.reshape(nz, nx, ny), 0, 2), 0, 1)
    return data.reshape(nz, ny, nx)


def read_image(path):
    """_Refl.dat / _Trans.dat / _FlR.dat / _FlT.dat: det.ny rows of det.nx
    values (row y, column x).  Returns an array of shape (ny, nx)."""
    rows = []
    with open(path) as f:
        for line in f:
            vals = line.split()
            if vals:
                rows.append([float(v) for v in vals])
    if not rows or any(len(r) != len(rows[0]) for r in rows):
        raise ValueError(f"{path}: ragged image")
    return np.array(rows, dtype=np.float64)


def read_time(path):
    """_time.dat: a comment line, then one line per temporal detector with its
    position and its histogram.  Returns (positions (n, 2), histograms (n, bins))."""
    pos, hist = [], []
    with open(path) as f:
        for line in f:
            if line.lstrip().startswith("#"):
                continue
            vals = line.split()
            if not vals:
                continue
            pos.append([float(vals[0]), float(vals[1])])
            hist.append([float(v) for v in vals[2:]])
    return np.array(pos, dtype=np.float64).reshape(-1, 2), np.array(hist, dtype=np.float64)
