"""Seeded voxel-medium generator and bulk-descriptor / .mci writers.

Shapes follow the reference's medium notebook (spheres and axis-parallel
cylinders, Bulk_descriptor_creator.ipynb:30-185); the bulk file is written in
its on-disk order, x fastest, one integer label per voxel separated by spaces
(Bulk_descriptor_creator.ipynb:249-254, read back by CUDAMCFLio.cu:618-636).
Optical ranges are the benchmark's, converted from 1/mm to the reference's
1/cm (mu[1/cm] = 10 mu[1/mm]).
"""
import numpy as np

from . import Simulation, SOURCE_PENCIL, SOURCE_DISK

# rows: n, mua, mus, flc, AbsCExi, AbsCEmi, eY, g (1/cm), .mci medium-line order


def random_materials(rng, n_mat=8):
    """Material table: mua U[0.05,1.0]/cm, mus U[50,200]/cm, g U[0.8,0.95],
    n = 1.4, fluorophore absorption U[0,0.2]/cm (flc = 1, AbsCEmi = AbsCExi),
    quantum yield 0.1."""
    rows = []
    for _ in range(n_mat):
        mua = rng.uniform(0.05, 1.0)
        mus = rng.uniform(50.0, 200.0)
        g = rng.uniform(0.8, 0.95)
        fl = rng.uniform(0.0, 0.2)
        rows.append([1.4, mua, mus, 1.0, fl, fl, 0.1, g])
    return np.asarray(rows, np.float32)


def inclusions(shape, rng, n_spheres=5, r_sphere=(3, 8), n_cyl=2, r_cyl=(2, 5), first_label=2):
    """uint8 labels (nz, ny, nx): background 1, then one new label per
    inclusion (spheres, then z-parallel cylinders), centres uniform inside."""
    nz, ny, nx = shape
    lab = np.ones(shape, np.uint8)
    z, y, x = np.ogrid[0:nz, 0:ny, 0:nx]
    label = first_label
    for _ in range(n_spheres):
        r = min(rng.uniform(*r_sphere), 0.5 * min(shape))
        c = rng.uniform([r, r, r], [nz - r, ny - r, nx - r])
        lab[(z + 0.5 - c[0]) ** 2 + (y + 0.5 - c[1]) ** 2 + (x + 0.5 - c[2]) ** 2 < r * r] = label
        label += 1
    for _ in range(n_cyl):
        r = min(rng.uniform(*r_cyl), 0.5 * min(ny, nx))
        c = rng.uniform([r, r], [ny - r, nx - r])
        mask = ((y + 0.5 - c[0]) ** 2 + (x + 0.5 - c[1]) ** 2 < r * r)
        lab[np.broadcast_to(mask, shape)] = label
        label += 1
    return lab


def inclusion_medium(seed=2024, shape=(64, 128, 128)):
    """128x128x64 medium of 0.5 mm voxels: 5 spheres r in [3,8] voxels, 2
    z-parallel cylinders r in [2,5] voxels, 8 materials."""
    rng = np.random.default_rng(seed)
    lab = inclusions(shape, rng)
    mats = random_materials(rng, 8)
    return lab, mats


def inclusion_simulation(photons=100_000_000, seed=2024, fhd=True, source=(0.0, 0.0, 0.0)):
    """The plain-path benchmark problem: grid_size 20 /cm (0.5 mm voxels),
    slab 3.2 cm, 128x128 whole-field images of 0.5 mm pixels, pencil beam."""
    lab, mats = inclusion_medium(seed)
    nz, ny, nx = lab.shape
    gs = 20
    return Simulation.voxel(lab, mats, esp=nz / gs, grid_size=gs, det=(1.0 / gs, 1.0 / gs, nx, ny, 0.0, 0.0),
                            photons=photons, source=source, src_kind=SOURCE_PENCIL, n_above=1.0, n_below=1.0,
                            fhd=fhd)


def sphere_simulation(photons=1_000_000, fhd=True, photons_per_voxel=40000):
    """The small plain problem in the reference's own grid shape: esp = 2 cm,
    grid_size 10 /cm, so the grid is (4 esp gs)^2 x (esp gs) = 80x80x20 voxels
    of 1 mm (CUDAMCFLio.cu:609-612); background mua 0.1, mus 100 /cm, g 0.9,
    n 1.4; a voxelized sphere of radius 0.5 cm centred at depth 1.0 cm with mua
    0.5 /cm; fluorophore absorption 0.01 /cm (background) and 0.1 /cm (sphere),
    quantum yield 0.1, emission absorption 0.08 /cm (AbsCEmi = (0.08 - mua) /
    flc); ambient n 1.0; 40x40 images of 1 mm pixels; pencil beam at the
    centre."""
    esp, gs = 2.0, 10
    nx = ny = int(4 * esp * gs)
    nz = int(esp * gs)
    z, y, x = np.ogrid[0:nz, 0:ny, 0:nx]
    # voxel centres in cm: x from -2 esp, z from 0
    xc = (x + 0.5) / gs - 2 * esp
    yc = (y + 0.5) / gs - 2 * esp
    zc = (z + 0.5) / gs
    lab = np.ones((nz, ny, nx), np.uint8)
    lab[xc * xc + yc * yc + (zc - 1.0) ** 2 < 0.5 * 0.5] = 2
    mats = [[1.4, 0.1, 100.0, 1.0, 0.01, 0.08 - 0.1, 0.1, 0.9],
            [1.4, 0.5, 100.0, 1.0, 0.1, 0.08 - 0.5, 0.1, 0.9]]
    return Simulation.voxel(lab, mats, esp=esp, grid_size=gs, det=(0.1, 0.1, 40, 40, 0.0, 0.0), photons=photons,
                            src_kind=SOURCE_PENCIL, n_above=1.0, n_below=1.0, fhd=fhd, reference_grid=True,
                            photons_per_voxel=photons_per_voxel)


def write_bulk(path, labels):
    """Bulk-descriptor file: labels x fastest, '%i' separated by spaces, one line."""
    flat = np.asarray(labels, np.uint8).reshape(-1)  # (nz, ny, nx) C-order == x fastest
    with open(path, "w") as f:
        f.write(" ".join(map(str, flat.tolist())))
        f.write("\n")


def write_mci(path, out_name, bulk_file, mats, esp, grid_size, photons, photons_per_voxel=40000, do_fl=1,
              det=(0.05, 0.05, 128, 128, 0.0, 0.0), source=(0.0, 0.0, 0.0), n_above=1.0, n_below=1.0, dims=None,
              temporal=None):
    """A method-2 .mci file in the reference grammar (CUDAMCFLio.cu:243-659).
    dims=(nx,ny,nz) appends the '# mcfl-dims' extension line after the bulk
    filename (a comment to the reference parser)."""
    dx, dy, nx, ny, x0, y0 = det
    t = temporal or (0.0, 0.0, 1, 1, 1.0, 1.0, 0.5, 0, 10, 1000000)
    lines = [
        "1.0 # file version",
        "1 # number of runs",
        "2 # bulk description method",
        f"{out_name} A # output file name, ASCII",
        f"{int(photons)} # No. of photons",
        f"{int(photons_per_voxel)} # No. of photons per voxel (fluorescence)",
        f"{int(grid_size)} # Grid finesse",
        f"{int(do_fl)} # Do fluorescence simulation?",
        f"{dx} {dy} # Image dx, dy",
        f"{nx} {ny} # Image No. of dx, dy",
        f"{x0} {y0} # Image x0, y0",
        f"{source[0]} {source[1]} {source[2]} # Source position x y z",
        f"{1 if temporal else 0} # Do temporal simulation?",
        f"{t[0]} {t[1]} # Detector array center x,y",
        f"{t[2]} {t[3]} # Number of detectors x,y",
        f"{t[4]} {t[5]} # Separation between detectors",
        f"{t[6]} # Temporal detector radius",
        f"{t[7]} # Location of the detector array",
        f"{t[8]} # Number of temporal bins",
        f"{t[9]} # Max temp [fs]",
        "",
        f"{len(mats)} # Number of bulk descriptors",
        f"{esp} # Bulk thickness",
        f"{n_above} # n for medium above",
    ]
    for r in mats:
        lines.append(" ".join(repr(float(v)) for v in r))
    lines.append(f"{n_below} # n for medium below")
    lines.append(bulk_file)
    if dims is not None:
        lines.append(f"# mcfl-dims {dims[0]} {dims[1]} {dims[2]}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


# ---------------------------------------------------------------- workloads

GRID_SIZES = [(64, 64, 64), (48, 96, 96), (64, 128, 128), (96, 192, 192), (128, 256, 256)]  # (nz, ny, nx)


def scan_sources(n=8, spacing_vox=8, gs=20):
    """8x8 pencil-beam positions spaced 8 voxels apart, centred (cm)."""
    step = spacing_vox / gs
    return [((i - (n - 1) / 2) * step, (j - (n - 1) / 2) * step, 0.0) for j in range(n) for i in range(n)]


def log_uniform(rng, lo, hi, size):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size)).astype(np.int64)


def scan_budgets(seed=7, n=64, lo=1e6, hi=5e7):
    return log_uniform(np.random.default_rng(seed), lo, hi, n)


def random_medium(rng, shape, gs=20):
    """One heterogeneous medium: 1-6 inclusions (spheres or z-cylinders),
    a 4-8 entry material table drawn from the benchmark ranges."""
    n_mat = int(rng.integers(4, 9))
    n_inc = int(rng.integers(1, 7))
    nz, ny, nx = shape
    n_sph = int(rng.integers(0, n_inc + 1))
    lab = inclusions(shape, rng, n_spheres=n_sph, n_cyl=n_inc - n_sph)
    # fold inclusion labels into the table size
    lab = np.where(lab > 1, 2 + (lab - 2) % (n_mat - 1), lab).astype(np.uint8)
    mats = random_materials(rng, n_mat)
    return Simulation.voxel(lab, mats, esp=nz / gs, grid_size=gs, det=(1.0 / gs, 1.0 / gs, nx, ny, 0.0, 0.0),
                            photons=1000, n_above=1.0, n_below=1.0, fhd=True)


def hetero_jobs(seed=11, n=32, lo=1e5, hi=1e8, gs=20):
    """32 independent simulations: grid sizes from GRID_SIZES, own media, a
    pencil or a uniform disk (radius 2-10 voxels) at a random lateral
    position, log-uniform budgets.  Returns (media, [(medium, kind, r, src, N, seed)])."""
    rng = np.random.default_rng(seed)
    media, jobs = [], []
    budgets = log_uniform(rng, lo, hi, n)
    for i in range(n):
        shape = GRID_SIZES[int(rng.integers(0, len(GRID_SIZES)))]
        media.append(random_medium(rng, shape, gs))
        nz, ny, nx = shape
        disk = bool(rng.integers(0, 2))
        r = float(rng.uniform(2, 10)) / gs if disk else 0.0
        src = (float(rng.uniform(-0.25, 0.25) * nx / gs), float(rng.uniform(-0.25, 0.25) * ny / gs), 0.0)
        jobs.append((i, SOURCE_DISK if disk else SOURCE_PENCIL, r, src, int(budgets[i]), 1000 + i))
    return media, jobs


def request_stream(seed=5, rate=3.0, duration=12.0, n_media=8, lo=1e5, hi=1e8, gs=20):
    """Poisson arrivals of forward simulations: media from the heterogeneous
    generator (restricted to grids up to 128x128x64) and scan-style sources,
    50% excitation-only / 50% excitation+fluorescence, log-uniform budgets.
    Returns (media, [(t_arrival, medium, src, N, seed, fluorescence)])."""
    rng = np.random.default_rng(seed)
    media = [random_medium(rng, GRID_SIZES[int(rng.integers(0, 3))], gs) for _ in range(n_media)]
    srcs = scan_sources(8, 8, gs)
    reqs, t = [], 0.0
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t > duration:
            break
        m = int(rng.integers(0, n_media))
        s = srcs[int(rng.integers(0, len(srcs)))]
        nz, ny, nx = media[m].s.nz, media[m].s.ny, media[m].s.nx
        s = (max(min(s[0], 0.45 * nx / gs), -0.45 * nx / gs), max(min(s[1], 0.45 * ny / gs), -0.45 * ny / gs), 0.0)
        n = int(log_uniform(rng, lo, hi, 1)[0])
        reqs.append((t, m, s, n, 5000 + len(reqs), bool(rng.integers(0, 2))))
    return media, reqs
