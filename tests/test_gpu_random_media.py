"""Randomised parity: 40 seeded random problems, each run on the GPU and by
the oracle, must give bit-identical integer tallies.  The problems vary
everything the path has: grid shape and voxel size, up to six materials
(glass, refractive-index steps, isotropic scattering, fluorophore
absorption), inclusions of random labels including the internal exit labels
0 and n_bulks+1, pencil, disk and isotropic sources anywhere in the slab,
detector windows that clip the exits, temporal detectors, every optional
tally, photon streams and the reference's lane streams.  The per-photon step
cap is 20000 on both sides, so a photon trapped in glass by total internal
reflection stops the same way on both; a photon launched inside an exit
label marches to the surface, where the march's pass bound ends it and it
leaves (the reference loops for ever there)."""
import numpy as np
import pytest

import cudamcfl_amd as M
import _oracle as O

pytestmark = pytest.mark.gpu

CAP = 20_000
A_TAB = M.multipliers(32768)


def random_problem(seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = (int(v) for v in rng.integers(1, [25, 25, 17]))
    gs = int(rng.choice([5, 10, 20]))
    n_mat = int(rng.integers(1, 7))
    mats = []
    for i in range(n_mat):
        n = float(rng.choice([1.33, 1.4, 1.5]))
        if i > 0 and rng.random() < 0.15:  # glass (not the background, which also fills the space beside the grid)
            mats.append([n, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, float(rng.uniform(0.0, 0.9))])
            continue
        g = 0.0 if rng.random() < 0.15 else float(rng.uniform(0.5, 0.95))
        fl = float(rng.uniform(0.0, 0.5)) if rng.random() < 0.7 else 0.0
        mats.append([n, float(rng.uniform(0.01, 3.0)), float(rng.uniform(5.0, 200.0)), 1.0, fl,
                     float(rng.uniform(0.0, 0.5)), 0.1, g])
    lab = np.ones((nz, ny, nx), np.uint8)
    for _ in range(int(rng.integers(0, 5))):  # boxes of random labels, now and then an internal exit
        z0, y0, x0 = (int(rng.integers(0, d)) for d in (nz, ny, nx))
        z1, y1, x1 = z0 + int(rng.integers(1, nz + 1)), y0 + int(rng.integers(1, ny + 1)), x0 + int(rng.integers(1, nx + 1))
        label = int(rng.integers(1, n_mat + 1)) if rng.random() < 0.9 else int(rng.choice([0, n_mat + 1]))
        lab[z0:z1, y0:y1, x0:x1] = label
    esp = nz / gs
    hx, hy = nx / (2 * gs), ny / (2 * gs)
    kind = int(rng.choice([M.SOURCE_PENCIL, M.SOURCE_DISK, M.SOURCE_ISOTROPIC]))
    src = (float(rng.uniform(-hx, hx)), float(rng.uniform(-hy, hy)),
           0.0 if (kind != M.SOURCE_ISOTROPIC and rng.random() < 0.6) else float(rng.uniform(0.0, esp * 0.99)))
    src_r = float(rng.uniform(0.01, max(hx, 0.02))) if kind == M.SOURCE_DISK else 0.0
    dnx, dny = int(rng.integers(1, 21)), int(rng.integers(1, 21))
    det = (float(rng.uniform(0.02, 0.2)), float(rng.uniform(0.02, 0.2)), dnx, dny,
           float(rng.uniform(-0.3, 0.3)), float(rng.uniform(-0.3, 0.3)))
    temporal = None
    if rng.random() < 0.5:
        temporal = (float(rng.uniform(-0.2, 0.2)), 0.0, int(rng.integers(1, 4)), int(rng.integers(1, 3)),
                    float(rng.uniform(0.05, 0.3)), float(rng.uniform(0.05, 0.3)), float(rng.uniform(0.05, 0.4)),
                    int(rng.integers(0, 2)), int(rng.integers(1, 60)), int(rng.integers(10_000, 5_000_000)))
    photons = int(rng.integers(5000, 40000))
    sim = M.Simulation.voxel(lab, mats, esp=esp, grid_size=gs, det=det, photons=photons, source=src,
                             src_kind=kind, src_r=src_r, n_above=float(rng.choice([1.0, 1.33])),
                             n_below=float(rng.choice([1.0, 1.4, 1.5])), fhd=bool(rng.random() < 0.8),
                             temporal=temporal)
    opts = dict(axyz=bool(rng.random() < 0.5), sq=bool(rng.random() < 0.5))
    lanes = rng.random() < 0.2
    return sim, opts, lanes, int(rng.integers(1, 1 << 30))


@pytest.fixture(scope="module")
def capped(request):
    mp = pytest.MonkeyPatch()
    mp.setenv("MCFL_MAX_STEPS", str(CAP))
    c = M.Context(0)
    mp.undo()
    O.set_max_steps(CAP)
    yield c
    O.set_max_steps(0)
    c.close()


@pytest.mark.parametrize("seed", range(40))
def test_random_problem_bit_exact(capped, seed):
    sim, opts, lanes, key = random_problem(1000 + seed)
    s = sim.s
    print(f"case: random problem {seed}: {s.nx}x{s.ny}x{s.nz} gs {s.grid_size}, {s.n_bulks} materials, "
          f"source {s.src_kind}/{s.dir}, {s.number_of_photons} photons, lanes={lanes}, {opts}", flush=True)
    o = O.FromProduct(sim, **opts)
    if lanes:
        x, a = M.init_rng(256, seed=key)
        xo = x.copy()
        g = capped.run_excitation(sim, lanes=(x, a), **opts)
        r = O.run_lanes(o.o, xo, a, s.number_of_photons)
        assert np.array_equal(x, xo), "advanced lane states"
    else:
        g = capped.run_excitation(sim, seed=key, **opts)
        r = O.run_photons(o.o, A_TAB, M.stream_key(key, 0), 0, s.number_of_photons)
    for name, oname in (("Rd", "Rd"), ("Tt", "Tt"), ("fhd", "fhd"), ("A_xyz", "axyz"), ("time", "time"),
                        ("Rd2", "Rd2"), ("Tt2", "Tt2")):
        a_, b_ = getattr(g, name, None), getattr(r, oname, None)
        if a_ is None or b_ is None:
            continue
        k = min(len(a_), len(b_))
        assert np.array_equal(a_[:k], b_[:k]), name
    for k in O.TOT_NAMES:
        assert g.totals[k] == r.tot[k], k
    t = g.totals
    assert t["launched_w"] + t["roul_gain"] - t["roul_loss"] - t["zero_kill"] - t["aborted"] == t["R"] + t["T"] + t["A"]
