"""GPU transport vs the CPU oracle on the same inputs: integer tallies must be
bit-identical (the kernel and the oracle execute the same FP32 arithmetic)."""
import numpy as np
import pytest

import cudamcfl_amd as M
import _oracle as O
from media import small_sim, small_labels, MATS3

pytestmark = pytest.mark.gpu

A_TAB = None


def a_tab():
    global A_TAB
    if A_TAB is None:
        A_TAB = M.multipliers(32768)
    return A_TAB


def assert_same(g, o, names=("Rd", "Tt", "fhd", "A_xyz", "time", "Rd2", "Tt2")):
    omap = {"A_xyz": "axyz"}
    for n in names:
        ga = getattr(g, n, None)
        oa = getattr(o, omap.get(n, n), None)
        if ga is None or oa is None:
            continue
        k = min(len(ga), len(oa))
        diff = np.nonzero(ga[:k] != oa[:k])[0]
        assert len(diff) == 0, f"{n}: {len(diff)} entries differ, first {diff[:5]} gpu {ga[diff[:5]]} cpu {oa[diff[:5]]}"


def check_totals(gt, ot):
    for k in O.TOT_NAMES:
        assert gt[k] == ot[k], f"total {k}: gpu {gt[k]} cpu {ot[k]}"


def balance(t):
    return t["launched_w"] + t["roul_gain"] - t["roul_loss"] - t["zero_kill"] - t["aborted"] - (t["R"] + t["T"] + t["A"])


def test_photon_streams_bit_exact(ctx):
    sim = small_sim(photons=20000)
    print("case: photon streams, 20x20x10, 20000 photons, seed 7", flush=True)
    g = ctx.run_excitation(sim, seed=7, axyz=True, sq=True)
    o = O.FromProduct(sim, axyz=True, sq=True)
    r = O.run_photons(o.o, a_tab(), M.stream_key(7, 0), 0, 20000)
    assert_same(g, r)
    check_totals(g.totals, r.tot)
    assert g.totals["photons"] == 20000
    assert balance(g.totals) == 0
    assert g.Rd.sum() > 0 and g.Tt.sum() > 0 and g.time.sum() > 0


@pytest.mark.parametrize("fhd,axyz", [(False, True), (True, False), (False, False)])
def test_volume_tally_kernels_bit_exact(ctx, fhd, axyz):
    """Launches compile only the volume tallies their jobs need (fhd and A_xyz,
    fhd alone, none): each of those kernels against the oracle, including
    A_xyz without fhd (which needs the full kernel)."""
    sim = small_sim(photons=20000, fhd=fhd)
    print(f"case: volume tallies fhd={fhd} axyz={axyz}, 20x20x10, 20000 photons, seed 8", flush=True)
    g = ctx.run_excitation(sim, seed=8, axyz=axyz, sq=True)
    o = O.FromProduct(sim, axyz=axyz, sq=True)
    r = O.run_photons(o.o, a_tab(), M.stream_key(8, 0), 0, 20000)
    assert_same(g, r)
    check_totals(g.totals, r.tot)
    assert (g.fhd is not None and g.fhd.sum() > 0) == fhd
    assert (g.A_xyz is not None and g.A_xyz.sum() > 0) == axyz


def test_lane_streams_bit_exact(ctx):
    sim = small_sim(photons=5003, temporal=False)
    x, a = M.init_rng(512, seed=12345)
    x_cpu = x.copy()
    print("case: reference lane streams, 512 lanes, 5003 photons", flush=True)
    g = ctx.run_excitation(sim, lanes=(x, a))
    o = O.FromProduct(sim)
    r = O.run_lanes(o.o, x_cpu, a, 5003)
    assert_same(g, r)
    check_totals(g.totals, r.tot)
    assert np.array_equal(x, x_cpu), "advanced lane states differ"
    assert g.totals["photons"] == 5003


def test_lanes_need_enough_photons(ctx):
    sim = small_sim(photons=100, temporal=False)
    x, a = M.init_rng(512, seed=3)
    with pytest.raises(M.McflError, match="too few photons"):
        ctx.run_excitation(sim, lanes=(x, a))


def test_fluorescence_bit_exact(ctx):
    lab = small_labels(10, 10, 5)
    sim = M.Simulation.voxel(lab, MATS3, esp=0.5, grid_size=10, det=(0.1, 0.1, 10, 10, 0.0, 0.0), photons=20000,
                             n_above=1.0, n_below=1.33, fhd=True)
    print("case: fluorescence stage, 10x10x5, 40 photons/voxel", flush=True)
    ex = ctx.run_excitation(sim, seed=11)
    Fx = ex.fhd.astype(np.float64) / (0xFFFFFFFF * ex.totals["photons"])
    fl = ctx.run_fluorescence(sim, Fx, 40, seed=99)
    vox, scale, unit = M.fl_plan(sim, Fx, 40)
    assert fl["voxels"] == len(vox) and fl["unit"] == unit
    keys = np.array([M.stream_key(99, int(v)) for v in vox], np.uint64)
    o = O.FromProduct(sim, emission=True)
    r = O.run_fluor_jobs(o.o, a_tab(), vox, keys, scale, 40)
    assert np.array_equal(fl["raw_R"], r.Rd) and np.array_equal(fl["raw_T"], r.Tt)
    check_totals(fl["totals"], r.tot)
    assert fl["raw_R"].sum() > 0
    dxdy = float(np.float32(0.1)) * float(np.float32(0.1))
    np.testing.assert_allclose(fl["FlR"], r.Rd[:100] * unit / dxdy, rtol=1e-12)


def test_glass_inclusion_bit_exact(ctx):
    lab = small_labels()
    lab[2:5, 5:12, 5:12] = 4  # a glass (mus = 0) block
    mats = MATS3 + [[1.45, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.9]]
    sim = M.Simulation.voxel(lab, mats, esp=1.0, grid_size=10, det=(0.1, 0.1, 20, 20, 0.0, 0.0), photons=8000,
                             n_above=1.0, n_below=1.0, fhd=True)
    print("case: glass inclusion, 8000 photons", flush=True)
    g = ctx.run_excitation(sim, seed=5, axyz=True)
    o = O.FromProduct(sim, axyz=True)
    r = O.run_photons(o.o, a_tab(), M.stream_key(5, 0), 0, 8000)
    assert_same(g, r)
    check_totals(g.totals, r.tot)


GOLDEN = __import__("pathlib").Path(__file__).parent / "golden" / "ref_inputs"

# three layers: a glass layer (mus = 0, the 2 cm step of :91/:99; mua = 0,
# since mua * FLT_MAX overflows the weight arithmetic in the reference), an
# index step inside the slab, and a scattering inclusion straddling two layers
LAYERS3 = [(1.4, 0.5, 80.0, 0.02, 5.0, 3.0, 0.1, 0.85, 0.05),
           (1.4, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.9, 0.02),
           (1.33, 1.0, 120.0, 0.0, 0.0, 0.0, 0.0, 0.9, 0.15)]
INC3 = (0.05, -0.02, 0.06, 0.05, 1.4, 2.0, 150.0, 0.1, 3.0, 2.0, 0.1, 0.7)


def layered3(photons):
    tmp = (0.0, 0.0, 2, 1, 0.3, 0.3, 0.25, 0, 40, 3000000)
    return M.Simulation.layered(LAYERS3, INC3, grid_size=20, det=(0.05, 0.05, 16, 16, 0.0, 0.0), photons=photons,
                                n_above=1.0, n_below=1.0, fhd=True, temporal=tmp)


def test_layered_photon_streams_bit_exact(ctx):
    sim = layered3(20000)
    print("case: layered slab (3 layers, glass, index step, inclusion), photon streams, 20000 photons", flush=True)
    g = ctx.run_excitation(sim, seed=31, axyz=True, sq=True)
    o = O.FromProduct(sim, axyz=True, sq=True)
    r = O.run_photons(o.o, a_tab(), M.stream_key(31, 0), 0, 20000)
    assert_same(g, r)
    check_totals(g.totals, r.tot)
    assert g.totals["photons"] == 20000 and g.totals["zero_kill"] == 0
    assert balance(g.totals) == 0
    assert g.Rd.sum() > 0 and g.Tt.sum() > 0 and g.fhd.sum() > 0 and g.time.sum() > 0


def test_layered_sample_mci_lanes_bit_exact(ctx):
    """The reference's own method-1 input on its 32768-lane geometry (reduced budget)."""
    sim = M.Simulation.from_mci(str(GOLDEN / "sample.mci"))
    sim.s.number_of_photons = 40000
    sim.s.fhd_activated = 1
    x, a = M.init_rng(M.NUM_THREADS, seed=777)
    x_cpu = x.copy()
    print("case: sample.mci (method 1), reference lanes, 40000 photons", flush=True)
    g = ctx.run_excitation(sim, lanes=(x, a))
    o = O.FromProduct(sim)
    r = O.run_lanes(o.o, x_cpu, a, 40000)
    assert_same(g, r)
    check_totals(g.totals, r.tot)
    assert np.array_equal(x, x_cpu), "advanced lane states differ"


def test_layered_batch_equals_single_runs(ctx):
    """Jobs on one layered medium (bulk method 1: three layers with glass, an
    index step, a straddling inclusion, temporal detectors) share one batched
    launch: every job's tallies equal its plain run alone, the starting layer
    taken from each job's own source depth; one job equals the oracle.  A
    layered medium next to another medium in one batch, an emission job on a
    layered medium (method 1 emits nothing) and a source outside every layer
    are refused."""
    sim = layered3(1000)
    srcs = [(0.0, 0.0, 0.0), (0.12, -0.05, 0.0), (-0.1, 0.08, 0.03)]
    budgets = [6000, 2500, 4000]
    tallies = M.TALLY_FHD | M.TALLY_SQ | M.TALLY_TEMP
    jobs = [M.BatchJob(0, n, 40 + i, source=src, tallies=tallies) for i, (src, n) in enumerate(zip(srcs, budgets))]
    print("case: batch of 3 sources on a layered medium", flush=True)
    b = M.Batch(ctx, [sim], jobs)
    b.run()
    b.close()
    for i, (src, n) in enumerate(zip(srcs, budgets)):
        sim.s.number_of_photons = n
        sim.s.xi, sim.s.yi, sim.s.zi = src
        r = ctx.run_excitation(sim, seed=40 + i, sq=True)
        j = jobs[i]
        for name in ("Rd", "Tt", "fhd", "time", "Rd2"):
            assert np.array_equal(getattr(j, name), getattr(r, name)), (i, name)
        assert j.totals == r.totals, i
    sim.s.xi, sim.s.yi, sim.s.zi = srcs[2]
    o = O.FromProduct(sim, sq=True)
    r = O.run_photons(o.o, a_tab(), M.stream_key(42, 0), 0, budgets[2])
    assert np.array_equal(jobs[2].Rd, r.Rd) and np.array_equal(jobs[2].fhd, r.fhd)
    check_totals(jobs[2].totals, r.tot)
    with pytest.raises(M.McflError):
        M.Batch(ctx, [sim, small_sim(1000)], [M.BatchJob(0, 1000, 1), M.BatchJob(1, 1000, 2)])
    with pytest.raises(M.McflError):
        M.Batch(ctx, [sim], [M.BatchJob(0, 1000, 1, src_kind=M.SOURCE_EMISSION, Fx=np.ones(sim.nvox))])
    with pytest.raises(M.McflError):
        M.Batch(ctx, [sim], [M.BatchJob(0, 1000, 1, source=(0.0, 0.0, 50.0))])


def test_zero_and_one_photon(ctx):
    """An empty run is refused as an argument error (the reference's parser
    never yields one); a single photon gives the oracle's tallies."""
    sim = small_sim(photons=0)
    print("case: 0 photons", flush=True)
    with pytest.raises(M.McflError):
        ctx.run_excitation(sim, seed=5, sq=True)
    print("case: 1 photon", flush=True)
    sim.s.number_of_photons = 1
    g = ctx.run_excitation(sim, seed=5, sq=True)
    o = O.FromProduct(sim, sq=True)
    r = O.run_photons(o.o, a_tab(), M.stream_key(5, 0), 0, 1)
    assert_same(g, r)
    check_totals(g.totals, r.tot)
    assert g.totals["photons"] == 1 and balance(g.totals) == 0


def test_single_voxel_medium(ctx):
    """A 1x1x1 grid: every lateral step leaves the grid (medium 1 outside,
    CUDAMCFLtransport.cu:319-333) and the fhd has one entry."""
    lab = np.ones((1, 1, 1), np.uint8)
    sim = M.Simulation.voxel(lab, MATS3[:1], esp=0.1, grid_size=10, det=(0.05, 0.05, 8, 8, 0.0, 0.0), photons=5000,
                             n_above=1.0, n_below=1.33, fhd=True)
    print("case: 1x1x1 medium, 5000 photons", flush=True)
    g = ctx.run_excitation(sim, seed=8, axyz=True)
    o = O.FromProduct(sim, axyz=True)
    r = O.run_photons(o.o, a_tab(), M.stream_key(8, 0), 0, 5000)
    assert_same(g, r)
    check_totals(g.totals, r.tot)
    assert g.fhd[0] > 0 and balance(g.totals) == 0


def test_largest_material_table(ctx):
    """98 materials (the reference's limit, labels 1..98 plus the two outer
    media) with varied n, g and a glass material: bit-exact vs the oracle."""
    rng = np.random.default_rng(17)
    nb = 98
    mats = []
    for i in range(nb):
        row = [float(rng.choice([1.33, 1.4, 1.45])), float(rng.uniform(0.05, 2.0)), float(rng.uniform(20, 120)),
               float(rng.uniform(0, 1)), 0.2, 0.1, 0.1, float(rng.uniform(0.0, 0.95))]
        if i == 40:  # glass (mus = 0, non-absorbing: the MoveToFirstBoundary path)
            row[1], row[2], row[3] = 0.0, 0.0, 0.0
        mats.append(row)
    lab = rng.integers(1, nb + 1, size=(8, 16, 16)).astype(np.uint8)
    sim = M.Simulation.voxel(lab, mats, esp=0.8, grid_size=10, det=(0.1, 0.1, 16, 16, 0.0, 0.0), photons=6000,
                             n_above=1.0, n_below=1.2, fhd=True)
    print("case: 98 materials, 16x16x8, 6000 photons", flush=True)
    g = ctx.run_excitation(sim, seed=9)
    o = O.FromProduct(sim)
    r = O.run_photons(o.o, a_tab(), M.stream_key(9, 0), 0, 6000)
    assert_same(g, r)
    check_totals(g.totals, r.tot)
    assert balance(g.totals) == 0


def test_reference_shaped_sphere_problem_bit_exact(ctx):
    """The small plain problem in the reference's own grid shape (80x80x20 of
    1 mm, a voxelized sphere): excitation and the every-voxel fluorescence
    stage bit-exact vs the oracle."""
    from cudamcfl_amd.media import sphere_simulation
    sim = sphere_simulation(photons=30000)
    print("case: 80x80x20 sphere problem, 30000 photons, seed 12", flush=True)
    g = ctx.run_excitation(sim, seed=12)
    o = O.FromProduct(sim)
    r = O.run_photons(o.o, a_tab(), M.stream_key(12, 0), 0, 30000)
    assert_same(g, r)
    check_totals(g.totals, r.tot)
    print("case: its fluorescence stage, 1 photon per emitting voxel", flush=True)
    Fx = g.fhd.astype(np.float64) / (0xFFFFFFFF * g.totals["photons"])
    fl = ctx.run_fluorescence(sim, Fx, 1, seed=13)
    vox, scale, unit = M.fl_plan(sim, Fx, 1)
    assert fl["voxels"] == len(vox) > 0
    keys = np.array([M.stream_key(13, int(v)) for v in vox], np.uint64)
    oe = O.FromProduct(sim, emission=True)
    re = O.run_fluor_jobs(oe.o, a_tab(), vox, keys, scale, 1)
    assert np.array_equal(fl["raw_R"], re.Rd) and np.array_equal(fl["raw_T"], re.Tt)
    check_totals(fl["totals"], re.tot)


def test_fluorescence_on_reference_lanes_bit_exact(ctx):
    """The emission loop on the reference's lane streams (voxels in the
    reference's order, lanes continuing from voxel to voxel, first photon of
    each lane replaying its launch draws): raw images, totals and the advanced
    lane states bit-exact vs the oracle; fewer photons per voxel than lanes is
    refused like the reference."""
    lab = small_labels(10, 10, 5, 4)
    sim = M.Simulation.voxel(lab, MATS3, esp=0.5, grid_size=10, det=(0.1, 0.1, 10, 10, 0.0, 0.0), photons=20000,
                             n_above=1.0, n_below=1.33, fhd=True)
    ex = ctx.run_excitation(sim, seed=11)
    Fx = ex.fhd.astype(np.float64) / (0xFFFFFFFF * ex.totals["photons"])
    L, nv = 512, 700
    x0, a = M.init_rng(L, seed=17)
    print("case: fluorescence on 512 reference lanes, 700 photons per voxel", flush=True)
    xg = x0.copy()
    fl = ctx.run_fluorescence_lanes(sim, Fx, nv, lanes=(xg, a))
    vox, scale, unit = M.fl_plan(sim, Fx, nv)
    vox, scale = O.reference_order(sim, vox, scale)
    xo = x0.copy()
    o = O.FromProduct(sim, emission=True)
    r = O.run_fluor_lanes(o.o, xo, a, vox, scale, nv)
    assert np.array_equal(fl["raw_R"], r.Rd) and np.array_equal(fl["raw_T"], r.Tt)
    check_totals(fl["totals"], r.tot)
    assert np.array_equal(xg, xo) and not np.array_equal(xg, x0)
    assert fl["raw_R"].sum() > 0 and fl["unit"] == unit
    with pytest.raises(M.McflError):
        ctx.run_fluorescence_lanes(sim, Fx, L - 1, lanes=(x0.copy(), a))


def test_per_voxel_emission_entry_point(ctx):
    """mcfl_DoOneSimulationFl (CUDAMCFLmain.cu:196-288) called once per
    voxel from a copy of the reference's voxel loop (:453-594: mua <- muaf,
    isotropic source at the voxel corner, 32768 lanes continuing from voxel
    to voxel): for three emitting voxels its raw R/T images, its photon count
    and the advanced lane states equal the oracle's lane run of the same
    voxel; fewer photons than lanes returns 0 like the reference."""
    lab = small_labels(10, 10, 5, 4)
    sim = M.Simulation.voxel(lab, MATS3, esp=0.5, grid_size=10, det=(0.1, 0.1, 10, 10, 0.0, 0.0), photons=40000,
                             n_above=1.0, n_below=1.33, fhd=False)
    s = sim.s
    for n in range(s.n_bulks + 2):
        s.bulks[n].mua = s.bulks[n].muaf
    s.dir, s.fhd_activated, s.do_temp_sim = 0.0, 0, 0
    nv = 40000
    s.number_of_photons = nv
    x0, a = M.init_rng(M.NUM_THREADS, seed=23)
    xg, xo = x0.copy(), x0.copy()
    o = O.FromProduct(sim)
    emitting = [v for v in range(lab.size) if s.bulks[int(lab.reshape(-1)[v])].albedof < 1]
    print("case: per-voxel emission entry point, 3 voxels of 40000 photons on 32768 lanes", flush=True)
    for v in emitting[:3]:
        iz, rem = divmod(v, 100)
        iy, ix = divmod(rem, 10)
        s.xi = np.float32(np.float32(ix) / np.float32(10)) - np.float32(s.halfx)
        s.yi = np.float32(np.float32(iy) / np.float32(10)) - np.float32(s.halfy)
        s.zi = np.float32(iz) / np.float32(10)
        n, R, T = ctx.DoOneSimulationFl(sim, xg, a)
        o.o.src_kind, o.o.xi, o.o.yi, o.o.zi = 0, s.xi, s.yi, s.zi
        r = O.run_lanes(o.o, xo, a, nv)
        assert n == nv == r.tot["photons"]
        assert np.array_equal(R, r.Rd) and np.array_equal(T, r.Tt) and R.sum() > 0
        assert np.array_equal(xg, xo)
    assert not np.array_equal(xg, x0)
    s.number_of_photons = M.NUM_THREADS - 1
    n, R, T = ctx.DoOneSimulationFl(sim, xg.copy(), a)
    assert n == 0 and R.sum() == 0


def test_step_cap_aborts_like_the_oracle(ctx, monkeypatch):
    """The per-photon step cap (the kernel's bounded-work watchdog) forced
    down to 40 steps from the host: photons still alive after 40 steps stop
    and their weight is tallied as aborted; tallies, totals and the energy
    balance equal the oracle run with the same cap."""
    monkeypatch.setenv("MCFL_MAX_STEPS", "40")
    capped = M.Context(0)
    sim = small_sim(photons=20000)
    print("case: photon streams, 20x20x10, step cap 40", flush=True)
    g = capped.run_excitation(sim, seed=31, axyz=True, sq=True)
    O.set_max_steps(40)
    try:
        o = O.FromProduct(sim, axyz=True, sq=True)
        r = O.run_photons(o.o, a_tab(), M.stream_key(31, 0), 0, 20000)
    finally:
        O.set_max_steps(0)
    assert_same(g, r)
    check_totals(g.totals, r.tot)
    assert g.totals["aborted"] > 0 and balance(g.totals) == 0
    assert g.totals["steps"] <= 40 * 20000
    capped.close()
