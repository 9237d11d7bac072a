"""The benchmark configurations at their real sizes with reduced photon
budgets: the 128x128x64 inclusion medium against the oracle bit for bit, the
64-source scan and the 32-medium heterogeneous batch (grids up to
256x256x128, disk sources) with every job equal to its run alone and some
against the oracle, and a served Poisson request stream with every request
equal to its work done alone."""
import numpy as np
import pytest

import cudamcfl_amd as M
import _oracle as O
from cudamcfl_amd.media import hetero_jobs, inclusion_simulation, request_stream, scan_budgets, scan_sources

pytestmark = pytest.mark.gpu

A_TAB = M.multipliers(32768)


def _same_tallies(g, r, names=("Rd", "Tt", "fhd", "Rd2", "Tt2")):
    omap = {"fhd": "fhd"}
    for n in names:
        a, b = getattr(g, n, None), getattr(r, omap.get(n, n), None)
        if a is None or b is None:
            continue
        assert np.array_equal(a, b), n


def test_bench_medium_bit_exact_against_the_oracle(ctx):
    """The 128x128x64 bench medium (8 materials, pencil beam, fhd on) at 2e5
    photons: images, fhd, squared weights and totals equal the oracle's.  fhd
    reaches voxels 20 and more voxels to the side of the source (beyond any
    LDS tile, flushed to the per-XCD copies) and the source column (inside
    the tile)."""
    N = 200_000
    sim = inclusion_simulation(photons=N)
    print("case: 128x128x64 bench medium, 2e5 photons vs the oracle", flush=True)
    g = ctx.run_excitation(sim, seed=33, sq=True)
    o = O.FromProduct(sim, sq=True)
    r = O.run_photons(o.o, A_TAB, M.stream_key(33, 0), 0, N)
    _same_tallies(g, r)
    for k in O.TOT_NAMES:
        assert g.totals[k] == r.tot[k], k
    fhd = g.fhd.reshape(64, 128, 128)
    assert fhd[:4, 60:68, 60:68].sum() > 0
    far = fhd[:, :, :44].sum() + fhd[:, :, 84:].sum()
    assert far > 0


def _scan_jobs(scale):
    srcs = scan_sources()
    budgets = np.maximum(1000, (scan_budgets() * scale).astype(np.int64))
    return srcs, budgets


def test_scan_batch_equals_plain_runs(ctx):
    """The 64-source scan (8x8 pencil beams 8 voxels apart on the bench
    medium, log-uniform budgets, here 1/1000 of the benchmark's) in one
    batched launch: every job's R/T images, squared weights and totals equal
    its plain run alone; two jobs equal the oracle."""
    sim = inclusion_simulation(photons=1000, fhd=False)
    srcs, budgets = _scan_jobs(1e-3)
    jobs = [M.BatchJob(0, int(n), 700 + i, source=s, tallies=M.TALLY_SQ) for i, (s, n) in enumerate(zip(srcs, budgets))]
    print(f"case: 64-source scan batch, {int(budgets.sum())} photons", flush=True)
    b = M.Batch(ctx, [sim], jobs)
    b.run()
    b.close()
    for i, (s, n) in enumerate(zip(srcs, budgets)):
        sim.s.number_of_photons = int(n)
        sim.s.xi, sim.s.yi, sim.s.zi = s
        r = ctx.run_excitation(sim, seed=700 + i, sq=True)
        j = jobs[i]
        assert np.array_equal(j.Rd, r.Rd) and np.array_equal(j.Tt, r.Tt) and np.array_equal(j.Rd2, r.Rd2), i
        assert j.totals == r.totals, i
    for i in (0, int(np.argmax(budgets))):
        sim.s.xi, sim.s.yi, sim.s.zi = srcs[i]
        o = O.FromProduct(sim, sq=True)
        r = O.run_photons(o.o, A_TAB, M.stream_key(700 + i, 0), 0, int(budgets[i]))
        assert np.array_equal(jobs[i].Rd, r.Rd) and np.array_equal(jobs[i].Tt, r.Tt)
        assert np.array_equal(jobs[i].Rd2, r.Rd2)
        for k in O.TOT_NAMES:
            assert jobs[i].totals[k] == r.tot[k], (i, k)


def test_heterogeneous_batch_equals_plain_runs(ctx):
    """32 independent media (grids 64^3 to 256x256x128, 1-6 inclusions, 4-8
    materials, pencil or uniform-disk sources, budgets 1/200 of the
    benchmark's) in one launch: every job equals its plain run alone; a disk
    job on a 256x256x128 grid equals the oracle."""
    media, spec = hetero_jobs()
    spec = [(m, kind, r, src, max(500, n // 200), seed) for (m, kind, r, src, n, seed) in spec]
    jobs = [M.BatchJob(m, n, seed, source=src, src_kind=kind, src_r=r) for (m, kind, r, src, n, seed) in spec]
    print(f"case: heterogeneous batch, 32 media, {sum(j[4] for j in spec)} photons", flush=True)
    b = M.Batch(ctx, media, jobs)
    b.run()
    b.close()
    for (m, kind, r, src, n, seed), j in zip(spec, jobs):
        s = media[m]
        s.s.number_of_photons, s.s.src_kind, s.s.src_r = n, kind, r
        s.s.xi, s.s.yi, s.s.zi = src
        s.s.fhd_activated = 0
        res = ctx.run_excitation(s, seed=seed)
        assert np.array_equal(j.Rd, res.Rd) and np.array_equal(j.Tt, res.Tt), m
        assert j.totals == res.totals, m
    k = next(i for i, (m, kind, r, src, n, seed) in enumerate(spec)
             if kind == M.SOURCE_DISK and media[m].s.nx == 256)
    m, kind, r, src, n, seed = spec[k]
    s = media[m]
    s.s.src_kind, s.s.src_r = kind, r
    s.s.xi, s.s.yi, s.s.zi = src
    o = O.FromProduct(s)
    res = O.run_photons(o.o, A_TAB, M.stream_key(seed, 0), 0, n)
    assert np.array_equal(jobs[k].Rd, res.Rd) and np.array_equal(jobs[k].Tt, res.Tt)
    for key in O.TOT_NAMES:
        assert jobs[k].totals[key] == res.tot[key], key


def test_served_request_stream_equals_alone(ctx):
    """A Poisson request stream (the benchmark's generator: media up to
    128x128x64, scan-style sources, half of the requests with fluorescence,
    budgets 1/1000 of the benchmark's), submitted in arrival order while the
    server steps: every request's excitation images and totals equal its
    plain run alone, and every fluorescence request's emission images equal
    an emission job on that run's fhd."""
    media, reqs = request_stream(duration=6.0)
    reqs = [(t, m, src, max(2000, n // 1000), seed, fl) for (t, m, src, n, seed, fl) in reqs]
    srv = M.Server(ctx, mem_budget=8 << 30, epoch_photons=5000)
    mids = [srv.add_medium(s) for s in media]
    print(f"case: served stream, {len(reqs)} requests", flush=True)
    ids, done, i, steps = {}, [], 0, 0
    while len(done) < len(reqs):
        if i < len(reqs):  # one arrival per step
            t, m, src, n, seed, fl = reqs[i]
            ids[srv.submit(mids[m], n, seed, source=src, fluorescence=fl)] = i
            i += 1
        d, _ = srv.step()
        done += d
        steps += 1
        assert steps < 5000
    assert sorted(done) == sorted(ids)
    n_fl = 0
    for rid, k in ids.items():
        t, m, src, n, seed, fl = reqs[k]
        out = srv.collect(rid, mids[m])
        s = media[m]
        s.s.number_of_photons = n
        s.s.xi, s.s.yi, s.s.zi = src
        s.s.src_kind = M.SOURCE_PENCIL
        s.s.fhd_activated = 1
        r = ctx.run_excitation(s, seed=seed)
        assert np.array_equal(out["Rd"], r.Rd) and np.array_equal(out["Tt"], r.Tt), k
        assert out["totals"] == r.totals, k
        if fl:
            n_fl += 1
            Fx = r.fhd.astype(np.float64) / (4294967295.0 * r.totals["photons"])
            job = M.BatchJob(0, n, seed, src_kind=M.SOURCE_EMISSION, Fx=Fx)
            b = M.Batch(ctx, [s], [job])
            b.run()
            b.close()
            assert np.array_equal(out["FlR"], job.Rd) and np.array_equal(out["FlT"], job.Tt), k
            assert out["fl_unit"] == job.fl_unit
    assert n_fl > 0
    srv.close()
