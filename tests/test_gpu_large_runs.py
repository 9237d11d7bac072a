"""Runs whose volume tallies approach the u64 range.  A plain photon-stream
excitation above a photon threshold runs as several launches of photons
[k0, k0 + n); photon k always draws from stream k, so the tallies equal those
of one launch.  A voxel whose summed weight exceeds 2^64 - 1 keeps 2^64 - 1
(the reference stops adding at LLONG_MAX, CUDAMCFLtransport.cu:364) instead of
wrapping, in one launch and across launches.  Contexts used from several host
threads at once give the results of the same runs made alone."""
import ctypes as C

import numpy as np
import pytest

import cudamcfl_amd as M
import _oracle as O
from cudamcfl_amd.media import inclusion_simulation
from media import small_sim

pytestmark = pytest.mark.gpu

U64_MAX = np.uint64(2 ** 64 - 1)


def _same(a, b):
    for n in ("Rd", "Tt", "fhd", "A_xyz", "time", "Rd2", "Tt2"):
        x, y = getattr(a, n, None), getattr(b, n, None)
        if x is None and y is None:
            continue
        assert np.array_equal(x, y), n
    assert a.totals == b.totals


def test_split_run_equals_one_launch(ctx, monkeypatch):
    """250000 photons as launches of 100000 + 100000 + 50000 (temporal
    detectors, fhd, A_xyz, squared weights): every tally equal to one launch's
    and to the oracle's."""
    sim = small_sim(photons=250000)
    monkeypatch.setenv("MCFL_RUN_PHOTONS", "100000")
    split = M.Context(0)
    print("case: 20x20x10, 250000 photons, one launch vs three", flush=True)
    a = ctx.run_excitation(sim, seed=9, axyz=True, sq=True)
    b = split.run_excitation(sim, seed=9, axyz=True, sq=True)
    _same(a, b)
    assert b.totals["photons"] == 250000 and b.fhd.sum() > 0 and b.time.sum() > 0
    o = O.FromProduct(sim, axyz=True, sq=True)
    r = O.run_photons(o.o, M.multipliers(32768), M.stream_key(9, 0), 0, 250000)
    assert np.array_equal(b.fhd, r.fhd) and np.array_equal(b.Rd, r.Rd)


def _hot_voxel_sim(photons):
    """A 3x3x1 grid of 1 cm voxels, isotropic scattering (mus 50/cm, g 0):
    each photon puts ~119 full weights of fhd into the source voxel, so 1e8
    photons need 2.8 times the u64 range there (and ~0.35 times in each of the
    8 per-XCD copies)."""
    lab = np.ones((1, 3, 3), np.uint8)
    return M.Simulation.voxel(lab, [(1.4, 0.01, 50.0, 0.0, 0.0, 0.0, 0.1, 0.0)], esp=1.0, grid_size=1,
                              det=(1.0, 1.0, 3, 3, 0.0, 0.0), photons=photons, n_above=1.0, n_below=1.0, fhd=True)


def test_hot_voxel_saturates_instead_of_wrapping(ctx, monkeypatch):
    N = 100_000_000
    small = ctx.run_excitation(_hot_voxel_sim(N // 8), seed=4)
    centre = int(np.argmax(small.fhd))
    assert float(small.fhd[centre]) * 8 > 2.0 ** 64 * 1.5  # the full run exceeds the u64 range there
    sim = _hot_voxel_sim(N)
    print("case: 3x3x1 hot voxel, 1e8 photons, one launch and four", flush=True)
    a = ctx.run_excitation(sim, seed=4, axyz=True)
    monkeypatch.setenv("MCFL_RUN_PHOTONS", "30000000")
    b = M.Context(0).run_excitation(sim, seed=4, axyz=True)
    for r in (a, b):
        assert r.totals["photons"] == N
        assert r.fhd[centre] == U64_MAX
        others = np.delete(r.fhd, centre)
        assert others.max() < U64_MAX and others.sum() > 0
    assert np.array_equal(np.delete(a.fhd, centre), np.delete(b.fhd, centre))
    assert np.array_equal(a.A_xyz, b.A_xyz) and np.array_equal(a.Rd, b.Rd) and np.array_equal(a.Tt, b.Tt)
    assert a.totals == b.totals


def test_contexts_on_two_host_threads(ctx):
    """The ABI is thread-compatible per context: two host threads, each with
    its own context (and stream), run excitations at the same time; each
    result equals the same run made alone."""
    import threading
    sims = [small_sim(photons=400000), _hot_voxel_sim(400000)]
    alone = [ctx.run_excitation(s, seed=21 + i, axyz=True) for i, s in enumerate(sims)]
    ctxs = [M.Context(0), M.Context(0)]
    out, errs = [None, None], []

    def work(i):
        try:
            for _ in range(3):
                out[i] = ctxs[i].run_excitation(sims[i], seed=21 + i, axyz=True)
        except Exception as e:  # reported below
            errs.append(e)

    print("case: two contexts on two host threads", flush=True)
    th = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not errs and all(not t.is_alive() for t in th)
    for a, b in zip(alone, out):
        _same(a, b)
    for c in ctxs:
        c.close()


def test_largest_grid_bit_exact():
    """A 1024x1024x512 grid (2^29 voxels: 512 MiB of labels, 4 GiB of fhd per
    copy, 32 GiB for the 8 per-XCD copies on the device): voxel indices, copy
    offsets and the copy reduction stay exact; fhd, images and totals equal
    the oracle's (one oracle thread, so it holds a single extra fhd array)."""
    from media import MATS3
    nz, ny, nx = 512, 1024, 1024
    lab = np.ones((nz, ny, nx), np.uint8)
    lab[20:60, 480:560, 480:560] = 2
    sim = M.Simulation.voxel(lab, MATS3, esp=nz / 20, grid_size=20, det=(0.05, 0.05, 256, 256, 0.0, 0.0),
                             photons=20000, fhd=True)
    del lab
    print("case: 1024x1024x512 grid, 20000 photons", flush=True)
    big = M.Context(0)
    g = big.run_excitation(sim, seed=13)
    big.close()
    o = O.FromProduct(sim)
    r = O.run_photons(o.o, M.multipliers(32768), M.stream_key(13, 0), 0, 20000, nthreads=1)
    assert g.totals == r.tot
    assert np.array_equal(g.Rd, r.Rd) and np.array_equal(g.Tt, r.Tt)
    assert g.fhd.sum() > 0 and np.array_equal(g.fhd, r.fhd)


def test_photon_ranges_add_up(ctx, monkeypatch):
    """Photons [0, 100000) and [100000, 300000) of a seed, run separately (the
    second also split into launches), add up to the run of [0, 300000): the
    primitive for splitting a run across GPUs or resuming it.  A photon
    offset with the reference's lane streams is refused."""
    N = 300_000
    full = ctx.run_excitation(small_sim(photons=N), seed=19, axyz=True, sq=True)
    a = ctx.run_excitation(small_sim(photons=100_000), seed=19, axyz=True, sq=True)
    monkeypatch.setenv("MCFL_RUN_PHOTONS", "70000")
    split = M.Context(0)
    print("case: photon ranges [0,1e5) + [1e5,3e5) vs [0,3e5)", flush=True)
    b = split.run_excitation(small_sim(photons=200_000), seed=19, axyz=True, sq=True, first_photon=100_000)
    split.close()
    for n in ("Rd", "Tt", "fhd", "A_xyz", "time", "Rd2", "Tt2"):
        assert np.array_equal(getattr(a, n) + getattr(b, n), getattr(full, n)), n
    for k in full.totals:
        assert a.totals[k] + b.totals[k] == full.totals[k], k
    x, am = M.init_rng(512, seed=3)
    rng = M.Rng()
    rng.mode, rng.x, rng.a, rng.n_lanes, rng.first_photon = M.RNG_LANES, x.ctypes.data_as(M.U64P), \
        am.ctypes.data_as(M.U32P), 512, 5
    t = M.Tallies()
    Rd = np.zeros(small_sim().npix, np.uint64)
    t.Rd_xy, t.Tt_xy = Rd.ctypes.data_as(M.U64P), Rd.ctypes.data_as(M.U64P)
    sim = small_sim(photons=5000)
    assert M.lib().mcfl_run_excitation(ctx.h, C.byref(sim.s), C.byref(rng), C.byref(t)) == 1  # MCFL_ERR_ARG


def test_split_run_on_the_bench_grid_equals_one_launch(ctx, monkeypatch):
    """The 128x128x64 bench medium at 4e6 photons run as four launches of 1e6
    (the per-launch copies summed into saturating accumulators, then copied
    back on the null stream after the context's stream has finished) gives
    the fhd, A_xyz, images and totals of one launch."""
    sim = inclusion_simulation(photons=4_000_000)
    a = ctx.run_excitation(sim, seed=17, axyz=True, sq=True)
    monkeypatch.setenv("MCFL_RUN_PHOTONS", "1000000")
    split = M.Context(0)
    print("case: 128x128x64, 4e6 photons, one launch vs four", flush=True)
    b = split.run_excitation(sim, seed=17, axyz=True, sq=True)
    split.close()
    _same(a, b)
    assert b.fhd.sum() > 0 and b.A_xyz.sum() > 0


def test_batch_and_server_do_not_wrap_a_hot_voxel(ctx):
    """4e8 photons on the 3x3x1 hot-voxel medium: one launch's 8 per-XCD fhd
    copies would each hold about 1.4 times the u64 range in the source voxel
    and wrap.  As a batch job (beside a small one) and as a served
    fluorescence request, the job runs in slices of 2^27 photons whose copies
    are summed into saturating accumulators: the batch job's fhd (2^64 - 1 in
    the source voxel), A_xyz, images and totals equal mcfl_run_excitation's
    split run, the small job its run alone, and the served request's
    excitation and emission images equal the split run and an emission job
    on its fhd."""
    N = 400_000_000
    sim = _hot_voxel_sim(N)
    print("case: 3x3x1 hot voxel, 4e8 photons, split run", flush=True)
    alone = ctx.run_excitation(sim, seed=4, axyz=True)
    centre = int(np.argmax(alone.fhd))
    assert alone.fhd[centre] == U64_MAX and np.delete(alone.fhd, centre).max() < U64_MAX
    jobs = [M.BatchJob(0, N, 4, tallies=M.TALLY_FHD | M.TALLY_AXYZ), M.BatchJob(0, 5000, 5, tallies=M.TALLY_FHD)]
    print("case: 3x3x1 hot voxel, 4e8 photons, batch in slices", flush=True)
    b = M.Batch(ctx, [sim], jobs)
    b.run()
    b.close()
    j = jobs[0]
    assert np.array_equal(j.fhd, alone.fhd) and np.array_equal(j.A_xyz, alone.A_xyz)
    assert np.array_equal(j.Rd, alone.Rd) and np.array_equal(j.Tt, alone.Tt)
    assert j.totals == alone.totals
    small = ctx.run_excitation(_hot_voxel_sim(5000), seed=5)
    assert np.array_equal(jobs[1].fhd, small.fhd) and np.array_equal(jobs[1].Rd, small.Rd)
    assert jobs[1].totals == small.totals
    print("case: 3x3x1 hot voxel, 4e8 photons, served fluorescence request", flush=True)
    srv = M.Server(ctx, mem_budget=1 << 30, epoch_photons=2_000_000)
    mid = srv.add_medium(sim)
    rid = srv.submit(mid, N, 4, fluorescence=True)
    done, steps = [], 0
    while not done:
        done, _ = srv.step()
        steps += 1
        assert steps < 2000
    out = srv.collect(rid, mid)
    srv.close()
    assert np.array_equal(out["Rd"], alone.Rd) and np.array_equal(out["Tt"], alone.Tt)
    assert out["totals"] == alone.totals
    Fx = alone.fhd.astype(np.float64) / (4294967295.0 * alone.totals["photons"])
    em = M.BatchJob(0, N, 4, src_kind=M.SOURCE_EMISSION, Fx=Fx)
    b = M.Batch(ctx, [sim], [em])
    b.run()
    b.close()
    assert np.array_equal(out["FlR"], em.Rd) and np.array_equal(out["FlT"], em.Tt)
    assert out["fl_unit"] == em.fl_unit
