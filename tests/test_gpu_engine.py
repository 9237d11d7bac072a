"""Batched launches and the continuous-batching server on the GPU.

Every job draws from photon streams keyed by its own seed, so a job's integer
tallies must be identical whether it runs alone, in a batch (one HIP-graph
captured launch), or sliced over serving epochs next to other requests."""
import numpy as np
import pytest

import cudamcfl_amd as M
from media import MATS3, small_labels

pytestmark = pytest.mark.gpu


def medium(nx=20, ny=20, nz=10, seed=0, n_below=1.33, temporal=False):
    lab = small_labels(nx, ny, nz, seed)
    tmp = (0.0, 0.0, 2, 1, 0.4, 0.4, 0.3, 0, 50, 1000000) if temporal else None
    return M.Simulation.voxel(lab, MATS3, esp=nz / 10, grid_size=10, det=(0.1, 0.1, nx, ny, 0.0, 0.0),
                              photons=1000, n_above=1.0, n_below=n_below, fhd=True, temporal=tmp)


def alone(ctx, sim, photons, seed, source):
    sim.s.number_of_photons = photons
    sim.s.xi, sim.s.yi, sim.s.zi = source
    return ctx.run_excitation(sim, seed=seed, sq=True)


def test_batch_equals_single_runs_and_graph_replays(ctx):
    sim = medium(temporal=True)
    srcs = [(0.0, 0.0, 0.0), (0.35, -0.2, 0.0), (-0.5, 0.4, 0.0)]
    budgets = [3000, 7000, 1500]
    jobs = [M.BatchJob(0, n, 100 + i, source=s, tallies=M.TALLY_FHD | M.TALLY_SQ | M.TALLY_TEMP)
            for i, (s, n) in enumerate(zip(srcs, budgets))]
    print("case: batch of 3 sources, graph replay x3", flush=True)
    b = M.Batch(ctx, [sim], jobs)
    first = []
    for rep in range(3):
        b.run()
        first.append([(j.Rd.copy(), j.fhd.copy(), j.time.copy()) for j in jobs])
        # eager work between replays
        alone(ctx, medium(), 500, 9, (0.0, 0.0, 0.0))
    for rep in (1, 2):
        for a, c in zip(first[0], first[rep]):
            assert all(np.array_equal(x, y) for x, y in zip(a, c)), f"replay {rep} differs"
    for i, j in enumerate(jobs):
        r = alone(ctx, sim, budgets[i], 100 + i, srcs[i])
        assert np.array_equal(j.Rd, r.Rd) and np.array_equal(j.Tt, r.Tt) and np.array_equal(j.fhd, r.fhd)
        assert np.array_equal(j.Rd2, r.Rd2) and np.array_equal(j.time, r.time)
        assert j.totals == r.totals
    # new seeds change the result, old seeds restore it
    b.run(seeds=[7, 8, 9])
    assert not np.array_equal(jobs[0].Rd, first[0][0][0])
    b.run(seeds=[100, 101, 102])
    assert np.array_equal(jobs[0].Rd, first[0][0][0])
    b.close()


def test_heterogeneous_batch_equals_single_runs(ctx):
    sims = [medium(20, 20, 10, 0), medium(12, 16, 6, 1, n_below=1.0), medium(24, 10, 8, 2)]
    jobs = [M.BatchJob(0, 4000, 1, tallies=M.TALLY_FHD), M.BatchJob(1, 2500, 2, source=(0.1, 0.1, 0.0)),
            M.BatchJob(2, 3000, 3, src_kind=M.SOURCE_DISK, src_r=0.3, tallies=M.TALLY_FHD),
            M.BatchJob(1, 1000, 4)]
    print("case: heterogeneous batch, 3 media, 4 jobs", flush=True)
    b = M.Batch(ctx, sims, jobs)
    b.run()
    for j in jobs:
        s = sims[j.medium]
        s.s.src_kind = j.src_kind
        s.s.src_r = j.src_r
        r = alone(ctx, s, j.photons, j.seed, j.source)
        s.s.src_kind = M.SOURCE_PENCIL
        assert np.array_equal(j.Rd, r.Rd) and np.array_equal(j.Tt, r.Tt)
        if j.fhd is not None:
            assert np.array_equal(j.fhd, r.fhd)
        assert j.totals == r.totals
    b.close()


def test_importance_sampled_emission_matches_voxel_loop(ctx):
    """Emission by alias sampling over voxels vs the reference's every-voxel
    loop: same FlR image within Monte Carlo noise."""
    sim = medium(10, 10, 5, 3)
    sim.s.esp = np.float32(0.5)
    ex = ctx.run_excitation(sim, seed=21)
    Fx = ex.fhd.astype(np.float64) / (0xFFFFFFFF * ex.totals["photons"])
    ref = ctx.run_fluorescence(sim, Fx, 2000, seed=5)
    n = ref["photons"]
    job = M.BatchJob(0, n, 77, src_kind=M.SOURCE_EMISSION, Fx=Fx, tallies=M.TALLY_SQ)
    b = M.Batch(ctx, [sim], [job])
    b.run()
    b.close()
    img = job.Rd[:100].astype(np.float64) * job.fl_unit / 0.01
    var = job.Rd2[:100].astype(np.float64) * 2.0 ** 32 * (job.fl_unit / 0.01) ** 2
    tot_ref, tot_is = ref["FlR"].sum(), img.sum()
    sig = np.sqrt(var.sum()) + 1e-300
    assert abs(tot_is - tot_ref) < 5 * sig + 0.02 * tot_ref, (tot_is, tot_ref, sig)


def test_server_results_equal_alone(ctx, monkeypatch):
    sims = [medium(20, 20, 10, 0), medium(12, 16, 6, 1, n_below=1.0)]
    monkeypatch.setenv("MCFL_EPOCH_TARGET", "4000")  # small epochs: every request is sliced over several
    srv = M.Server(ctx, mem_budget=1 << 30, epoch_photons=3000)
    mids = [srv.add_medium(s) for s in sims]
    reqs = [(mids[0], 5000, 11, (0.0, 0.0, 0.0), False), (mids[1], 2000, 12, (0.1, 0.0, 0.0), True),
            (mids[0], 800, 13, (-0.3, 0.2, 0.0), False), (mids[1], 6000, 14, (0.0, 0.1, 0.0), True)]
    ids = {}
    print("case: server, 4 requests over 2 media, epoch 3000 photons", flush=True)
    for k, (m, n, seed, src, fl) in enumerate(reqs[:2]):
        ids[srv.submit(m, n, seed, source=src, fluorescence=fl)] = k
    done = []
    steps = 0
    while len(done) < len(reqs):
        if steps == 2:
            for k, (m, n, seed, src, fl) in enumerate(reqs[2:], start=2):
                ids[srv.submit(m, n, seed, source=src, fluorescence=fl)] = k
        d, _ = srv.step()
        done += d
        steps += 1
        assert steps < 200
    assert srv.stats() == {"queued": 0, "active": 0, "bytes": 0}
    for rid, k in ids.items():
        m, n, seed, src, fl = reqs[k]
        out = srv.collect(rid, m)
        sim = sims[m]
        sim.s.number_of_photons = n
        sim.s.xi, sim.s.yi, sim.s.zi = src
        r = ctx.run_excitation(sim, seed=seed)
        assert np.array_equal(out["Rd"], r.Rd) and np.array_equal(out["Tt"], r.Tt), k
        assert out["totals"] == r.totals
        if n > 3000:
            assert out["epochs"] > (2 if fl else 1)
        if fl:
            assert out["epochs"] > 1
            Fx = r.fhd.astype(np.float64) / (4294967295.0 * r.totals["photons"])
            job = M.BatchJob(0, n, seed, src_kind=M.SOURCE_EMISSION, Fx=Fx)
            b = M.Batch(ctx, [sim], [job])
            b.run()
            b.close()
            assert np.array_equal(out["FlR"], job.Rd) and np.array_equal(out["FlT"], job.Tt), k
            assert out["fl_unit"] == job.fl_unit
    srv.close()


def test_server_admission_by_footprint(ctx):
    sim = medium(20, 20, 10, 0)
    one = 2 * sim.npix * 8 + 16 * 8  # footprint of an excitation-only request: R and T images + totals
    srv = M.Server(ctx, mem_budget=one + one // 2, epoch_photons=2000)  # room for one request at a time
    mid = srv.add_medium(sim)
    ids = [srv.submit(mid, 1500, s) for s in range(3)]
    done, _ = srv.step()
    st = srv.stats()
    assert st["active"] <= 1 and st["queued"] >= 1
    for _ in range(50):
        d, _ = srv.step()
        done += d
        if len(done) == 3:
            break
    assert sorted(done) == sorted(ids)
    # one request per epoch: the epochs run single-job launches, whose results
    # must still equal the request run alone
    sim.s.number_of_photons = 1500
    sim.s.xi, sim.s.yi, sim.s.zi = 0.0, 0.0, 0.0  # the requests' source
    for seed, rid in enumerate(ids):
        out = srv.collect(rid, mid)
        r = ctx.run_excitation(sim, seed=seed)
        assert np.array_equal(out["Rd"], r.Rd) and np.array_equal(out["Tt"], r.Tt), seed
        assert out["totals"] == r.totals, seed
    with pytest.raises(M.McflError):
        srv.submit(mid, 100, 1, fluorescence=True)  # its footprint exceeds the whole budget
    srv.close()


def test_autotuned_batch_equals_default(ctx):
    """Autotuning changes launch settings only: the replayed batch gives the
    same integer tallies as before tuning."""
    sim = medium(temporal=True)
    jobs = [M.BatchJob(0, 6000, 50 + i, source=s, tallies=M.TALLY_FHD | M.TALLY_SQ | M.TALLY_TEMP)
            for i, s in enumerate([(0.0, 0.0, 0.0), (0.3, 0.1, 0.0)])]
    print("case: batch of 2 sources, autotune at fraction 1/4", flush=True)
    b = M.Batch(ctx, [sim], jobs)
    b.run()
    before = [(j.Rd.copy(), j.fhd.copy(), j.time.copy(), dict(j.totals)) for j in jobs]
    t = b.autotune(0.25)
    assert t["trials"] > 1 and t["refill_min"] in (4, 8) and t["chunk_max"] in (256, 1024)
    assert t["wg"] in (512, 768, 1024)
    assert t["tile_side"] > 0 and t["trial_ms"] <= t["default_ms"]
    b.run()
    for j, (rd, fhd, tim, tot) in zip(jobs, before):
        assert np.array_equal(j.Rd, rd) and np.array_equal(j.fhd, fhd) and np.array_equal(j.time, tim)
        assert j.totals == tot
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("mixed", [False, True])
def test_batch_settings_do_not_change_results(ctx, mixed):
    """Every workgroup size, tile side, refill threshold and chunk cap the
    autotuner can choose gives the same integer tallies (one medium, and two
    media in one launch)."""
    sims = [medium(temporal=True)] + ([medium(16, 12, 8, 1, temporal=True)] if mixed else [])
    srcs = [(0.0, 0.0, 0.0), (0.3, 0.1, 0.0), (-0.2, 0.2, 0.0)]
    jobs = [M.BatchJob((i % len(sims)), 5000 + 700 * i, 70 + i, source=s,
                       tallies=M.TALLY_FHD | M.TALLY_SQ | M.TALLY_TEMP) for i, s in enumerate(srcs)]
    b = M.Batch(ctx, sims, jobs)
    print(f"case: settings sweep, mixed={mixed}", flush=True)
    b.run()
    ref = [(j.Rd.copy(), j.Tt.copy(), j.fhd.copy(), j.time.copy(), j.Rd2.copy(), dict(j.totals)) for j in jobs]
    for wg, side, refill, chunk in [(512, 0, 4, 1024), (1024, 0, 8, 256), (768, 16, 1, 64), (1024, 32, 64, 4096)]:
        print(f"case: wg={wg} side={side} refill={refill} chunk={chunk}", flush=True)
        b.configure(wg=wg, tile_side=side, refill_min=refill, chunk_max=chunk)
        b.run()
        for j, (rd, tt, fhd, tim, rd2, tot) in zip(jobs, ref):
            assert np.array_equal(j.Rd, rd) and np.array_equal(j.Tt, tt) and np.array_equal(j.fhd, fhd)
            assert np.array_equal(j.time, tim) and np.array_equal(j.Rd2, rd2) and j.totals == tot
    with pytest.raises(M.McflError):
        b.configure(wg=640)
    b.close()
