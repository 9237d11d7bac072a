"""Benchmark of the plain transport path on one MI355X.

Workload: the 128x128x64-voxel seeded medium (0.5 mm voxels, 5 spheres + 2
cylinders, 8 materials), pencil beam at the centre, 1e8 photons per step,
photon hitting density (3D) and 128x128 reflectance/transmittance images
tallied.  One step = one whole simulation = one persistent kernel launch.
`value` is photons per second of device time with the medium resident in HBM
(HIP events on the launch stream); the wall time including host<->device
copies is reported beside it.

Prints one JSON line.  --gpus N>1 exits: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

F_STEP = 150            # VALU issue slots per photon-step of the reference algorithm
PEAK_SLOTS = 256 * 4 * 16 * 2.4e9   # MI355X unpacked FP32 lane-issue rate (slots/s)


def cpu_baseline(target_s=20.0):
    """The CPU oracle on host cores, photon streams, same medium; a bounded
    sample sized to about target_s seconds."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import cudamcfl_amd as M
    import _oracle as O
    from cudamcfl_amd.media import inclusion_simulation

    threads = min(16, os.cpu_count() or 1)
    sim = inclusion_simulation(photons=10_000)
    o = O.FromProduct(sim)
    a_tab = M.multipliers(32768)
    key = M.stream_key(99, 0)
    n = 4000
    for _ in range(3):  # calibrate the sample size on growing runs (thread start-up dominates tiny ones)
        t0 = time.perf_counter()
        r = O.run_photons(o.o, a_tab, key, 0, n, nthreads=threads)
        dt = time.perf_counter() - t0
        if dt > 0.5:
            break
        n *= 8
    n = int(max(4000, min(40_000_000, n * target_s / max(dt, 1e-3))))
    t0 = time.perf_counter()
    r = O.run_photons(o.o, a_tab, key, 0, n, nthreads=threads)
    dt = time.perf_counter() - t0
    # one host thread, a ~3 s sample
    n1 = max(2000, int(n / threads * 3.0 / target_s))
    t0 = time.perf_counter()
    r1 = O.run_photons(o.o, a_tab, key, 0, n1, nthreads=1)
    dt1 = time.perf_counter() - t0
    return {"value": n / dt, "unit": "photons/s", "cores": threads, "kind": "port",
            "sample": f"{n} photons of the same 128x128x64 medium (photon streams), {dt:.1f} s",
            "photon_steps_per_s": r.tot["steps"] / dt,
            "single_thread": {"photons_per_s": n1 / dt1, "photon_steps_per_s": r1.tot["steps"] / dt1,
                              "sample": f"{n1} photons, {dt1:.1f} s"}}


def measured_traffic(photons):
    """HBM bytes per launch of this workload from the committed counter
    profile (FETCH_SIZE and WRITE_SIZE, each scaled on a plain copy); None when
    the profile covers another photon count."""
    path = os.path.join(ROOT, "profiles", "mcfl_transport_kernel_128x128x64_1e8_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    return t["bytes_per_launch"] if t.get("photons") == photons else None


def copy_probe(nbytes):
    """A plain device-to-device copy of known size: the calibration point of
    the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def rse_stats(jobs):
    """Per-pixel relative standard error of the reflectance images from the
    sum and sum-of-squares tallies: RSE = sqrt(Q - S^2/N) / S, over pixels
    above 1e-4 of each image's peak."""
    vals = []
    for j in jobs:
        S = j.Rd.astype(np.float64)
        Q = j.Rd2.astype(np.float64) * 2.0 ** 32
        keep = S > 1e-4 * S.max()
        r = np.sqrt(np.maximum(Q[keep] - S[keep] ** 2 / j.photons, 0.0)) / S[keep]
        vals.append(r)
    v = np.concatenate(vals)
    return {"median": float(np.median(v)), "p90": float(np.percentile(v, 90)),
            "frac_le_1pct": float(np.mean(v <= 0.01)), "pixels": int(v.size)}


def bench_scan(ctx):
    """64-source scan on the 128x128x64 medium: one graph-captured batched
    launch vs 64 plain launches with the same budgets and seeds (bit-identical
    images, hence identical per-pixel RSE)."""
    import cudamcfl_amd as M
    from cudamcfl_amd.media import inclusion_simulation, scan_sources, scan_budgets
    sim = inclusion_simulation(photons=1000, fhd=False)
    srcs, budgets = scan_sources(), scan_budgets()
    total = int(sum(budgets))
    jobs = [M.BatchJob(0, int(n), 700 + i, source=s, tallies=M.TALLY_SQ) for i, (s, n) in enumerate(zip(srcs, budgets))]
    b = M.Batch(ctx, [sim], jobs)
    b.run()
    t0 = time.perf_counter()
    tune = b.autotune(1.0 / 16)
    tune["wall_s"] = time.perf_counter() - t0
    print(f"scan autotune: {tune}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    kms = b.run()
    wall_b = time.perf_counter() - t0
    print(f"scan batched: {wall_b:.2f} s wall, {kms:.0f} ms device", file=sys.stderr, flush=True)
    plain_wall, plain_ms, same = 0.0, 0.0, True
    for i, (s, n) in enumerate(zip(srcs, budgets)):
        sim.s.number_of_photons = int(n)
        sim.s.xi, sim.s.yi, sim.s.zi = s
        t0 = time.perf_counter()
        r = ctx.run_excitation(sim, seed=700 + i, sq=True)
        plain_wall += time.perf_counter() - t0
        plain_ms += r.kernel_ms
        same &= bool(np.array_equal(r.Rd, jobs[i].Rd) and np.array_equal(r.Rd2, jobs[i].Rd2))
    print(f"scan plain: {plain_wall:.2f} s wall, {plain_ms:.0f} ms device, identical={same}", file=sys.stderr,
          flush=True)
    # the reference's launch geometry: 32768 MWC lanes per simulation (DoOneSimulation)
    lanes_wall, lanes_ms, lane_res = 0.0, 0.0, []
    for i, (s, n) in enumerate(zip(srcs, budgets)):
        sim.s.number_of_photons = int(n)
        sim.s.xi, sim.s.yi, sim.s.zi = s
        x, a = M.init_rng(M.NUM_THREADS, seed=900 + i)
        t0 = time.perf_counter()
        r = ctx.run_excitation(sim, lanes=(x, a), sq=True)
        lanes_wall += time.perf_counter() - t0
        lanes_ms += r.kernel_ms
        r.photons = int(n)
        lane_res.append(r)
    print(f"scan plain, reference lanes: {lanes_wall:.2f} s wall, {lanes_ms:.0f} ms device", file=sys.stderr,
          flush=True)
    out = {"workload": "8x8 pencil scan, 8-voxel pitch, budgets log-uniform [1e6,5e7] (seed 7), 128x128x64 medium",
           "photons": total, "batched_photons_per_s": total / wall_b, "batched_device_ms": kms,
           "plain_photons_per_s": total / plain_wall, "plain_device_ms": plain_ms,
           "speedup_wall": plain_wall / wall_b, "speedup_device": plain_ms / kms,
           "images_identical": same, "rse": rse_stats(jobs), "autotune": tune,
           "plain_reference_lanes": {"lanes": M.NUM_THREADS, "photons_per_s": total / lanes_wall,
                                     "device_ms": lanes_ms, "speedup_wall": lanes_wall / wall_b,
                                     "rse": rse_stats(lane_res)}}
    b.close()
    return out


def bench_hetero(ctx):
    """32 heterogeneous simulations (5 grid sizes up to 256x256x128, own media,
    pencil or disk sources, budgets log-uniform [1e5,1e8]) in one launch vs one
    launch each."""
    import cudamcfl_amd as M
    from cudamcfl_amd.media import hetero_jobs
    media, spec = hetero_jobs()
    total = sum(j[4] for j in spec)
    jobs = [M.BatchJob(m, n, seed, source=src, src_kind=kind, src_r=r) for (m, kind, r, src, n, seed) in spec]
    b = M.Batch(ctx, media, jobs)
    t0 = time.perf_counter()
    tune = b.autotune(1.0 / 16)
    tune["wall_s"] = time.perf_counter() - t0
    print(f"hetero autotune: {tune}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    kms = b.run()
    wall_b = time.perf_counter() - t0
    plain_wall, plain_ms, same = 0.0, 0.0, True
    for (m, kind, r, src, n, seed), j in zip(spec, jobs):
        s = media[m]
        s.s.number_of_photons, s.s.src_kind, s.s.src_r = n, kind, r
        s.s.xi, s.s.yi, s.s.zi = src
        s.s.fhd_activated = 0
        t0 = time.perf_counter()
        res = ctx.run_excitation(s, seed=seed)
        plain_wall += time.perf_counter() - t0
        plain_ms += res.kernel_ms
        same &= bool(np.array_equal(res.Rd, j.Rd))
    b.close()
    return {"workload": "32 media, grids 64^3..256x256x128, 1-6 inclusions, 4-8 materials, pencil/disk, "
                        "budgets log-uniform [1e5,1e8] (seed 11)", "photons": int(total),
            "batched_photons_per_s": total / wall_b, "batched_device_ms": kms,
            "plain_photons_per_s": total / plain_wall, "plain_device_ms": plain_ms,
            "speedup_wall": plain_wall / wall_b, "images_identical": same, "autotune": tune}


def _served(ctx, media, reqs, epoch):
    """Continuous batching: requests are submitted at their arrival times and
    the server steps whenever it has work.  Returns latencies, span, counters."""
    import cudamcfl_amd as M
    srv = M.Server(ctx, mem_budget=48 << 30, epoch_photons=epoch)
    mids = [srv.add_medium(m) for m in media]
    t0 = time.perf_counter()
    arr, lat, i = {}, {}, 0
    while i < len(reqs) or len(lat) < len(reqs):
        now = time.perf_counter() - t0
        while i < len(reqs) and reqs[i][0] <= now:
            ta, m, src, n, seed, fl = reqs[i]
            rid = srv.submit(mids[m], n, seed, source=src, fluorescence=fl)
            arr[rid] = (ta, mids[m])
            i += 1
        st = srv.stats()
        if st["active"] == 0 and st["queued"] == 0:
            if i < len(reqs):
                time.sleep(max(0.0, reqs[i][0] - (time.perf_counter() - t0)))
            continue
        done, _ = srv.step()
        tnow = time.perf_counter() - t0
        for rid in done:
            lat[rid] = tnow - arr[rid][0]
            srv.collect(rid, arr[rid][1])
    span = time.perf_counter() - t0
    cnt = srv.counters()
    srv.close()
    return np.array(sorted(lat.values())), span, cnt


def _plain(ctx, media, reqs, lanes=False):
    """One launch per request (+ one emission launch) in arrival order; with
    lanes, on the reference's 32768 per-lane MWC streams."""
    import cudamcfl_amd as M
    t0 = time.perf_counter()
    lat = []
    for k, (ta, m, src, n, seed, fl) in enumerate(reqs):
        now = time.perf_counter() - t0
        if now < ta:
            time.sleep(ta - now)
        s = media[m]
        s.s.number_of_photons = max(n, M.NUM_THREADS) if lanes else n
        s.s.xi, s.s.yi, s.s.zi = src
        s.s.src_kind = M.SOURCE_PENCIL
        s.s.fhd_activated = 1 if fl else 0
        if lanes:
            x, a = M.init_rng(M.NUM_THREADS, seed=seed)
            r = ctx.run_excitation(s, lanes=(x, a))
        else:
            r = ctx.run_excitation(s, seed=seed)
        if fl:
            Fx = r.fhd.astype(np.float64) / (4294967295.0 * r.totals["photons"])
            job = M.BatchJob(0, n, seed, src_kind=M.SOURCE_EMISSION, Fx=Fx)
            bb = M.Batch(ctx, [s], [job])
            bb.run()
            bb.close()
        lat.append(time.perf_counter() - t0 - ta)
        if lanes and k % 8 == 7:
            print(f"plain, reference lanes: {k + 1} requests", file=sys.stderr, flush=True)
    return np.array(sorted(lat)), time.perf_counter() - t0


def bench_serve(ctx, rate=3.0, high_rate=12.0, duration=12.0, epoch=250_000):
    """Poisson request streams: continuous batching vs a plain
    one-launch-per-request FIFO loop on the same arrivals, at a light load
    (also against the reference's launch geometry) and at a high load."""
    from cudamcfl_amd.media import request_stream

    def lat_stats(lat, work, span):
        return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
                "photons_per_s": work / span, "span_s": span}

    def one(r, with_lanes):
        media, reqs = request_stream(rate=r, duration=duration)
        work = sum(n * (2 if fl else 1) for (_, _, _, n, _, fl) in reqs)
        served, span_s, cnt = _served(ctx, media, reqs, epoch)
        print(f"served at {r}/s: {len(served)} requests, span {span_s:.1f} s", file=sys.stderr, flush=True)
        plain, span_p = _plain(ctx, media, reqs)
        print(f"plain at {r}/s: span {span_p:.1f} s", file=sys.stderr, flush=True)
        out = {"workload": f"Poisson arrivals {r}/s for {duration}s ({len(reqs)} requests), 8 media up to "
                           f"128x128x64, scan sources, 50% with fluorescence, budgets log-uniform [1e5,1e8]",
               "photons": int(work), "epoch_photons": epoch,
               "served": dict(lat_stats(served, work, span_s), epochs=cnt["epochs"],
                              device_photons_per_s=cnt["photons"] / max(cnt["kernel_ms"], 1e-9) * 1e3,
                              kernel_s=cnt["kernel_ms"] / 1e3, host_s=cnt["host_ms"] / 1e3),
               "plain": lat_stats(plain, work, span_p)}
        if with_lanes:
            lanes, span_l = _plain(ctx, media, reqs, lanes=True)
            print(f"plain, reference lanes: span {span_l:.1f} s", file=sys.stderr, flush=True)
            out["plain_reference_lanes"] = lat_stats(lanes, work, span_l)
        return out

    out = one(rate, True)
    out["high_load"] = one(high_rate, False)
    return out


def bench_sample2(ctx, frac=1.0):
    """The reference's own sample2.mci end to end: excitation (temporal
    detectors on) and the fluorescence stage over every emitting voxel,
    photon counts scaled by frac."""
    import numpy as np
    import cudamcfl_amd as M
    d = os.path.join(ROOT, "tests", "golden", "ref_inputs")
    cwd = os.getcwd()
    os.chdir(d)  # the bulk file is named relative to the working directory
    try:
        sim = M.Simulation.from_mci("sample2.mci")
    finally:
        os.chdir(cwd)
    s = sim.s
    n_ex = max(1, int(s.number_of_photons * frac))
    n_v = max(1, int(s.number_of_photons_per_voxel * frac))
    s.number_of_photons = n_ex
    t0 = time.perf_counter()
    ex = ctx.run_excitation(sim, seed=1)
    t1 = time.perf_counter()
    Fx = ex.fhd.astype(np.float64) / (4294967295.0 * ex.totals["photons"])
    fl = ctx.run_fluorescence(sim, Fx, n_v, seed=2)
    t2 = time.perf_counter()
    # the same images by importance-sampled emission (the CLI's -I): a tenth
    # of the every-voxel budget
    n_is = max(1, int(fl["photons"]) // 10)
    job = M.BatchJob(0, n_is, 3, src_kind=M.SOURCE_EMISSION, Fx=Fx, tallies=M.TALLY_SQ)
    b = M.Batch(ctx, [sim], [job])
    t3 = time.perf_counter()
    kms_is = b.run()
    t4 = time.perf_counter()
    b.close()
    xy = s.det.nx * s.det.ny
    k = job.fl_unit / (float(np.float32(s.det.dx)) * float(np.float32(s.det.dy)))
    is_tot = float(job.Rd[:xy].astype(np.float64).sum()) * k
    is_sig = float(np.sqrt(job.Rd2[:xy].astype(np.float64).sum() * 2.0 ** 32)) * k
    return {"workload": f"sample2.mci ({s.nx}x{s.ny}x{s.nz} voxels from 3dbulk-sample, temporal detectors), "
                        f"photon counts x{frac:g}",
            "excitation": {"photons": n_ex, "device_ms": ex.kernel_ms, "wall_s": t1 - t0,
                           "photons_per_s": n_ex / ex.kernel_ms * 1e3,
                           "steps_per_photon": ex.totals["steps"] / n_ex},
            "fluorescence": {"voxels": fl["voxels"], "photons_per_voxel": n_v, "photons": fl["photons"],
                             "device_ms": fl["kernel_ms"], "wall_s": t2 - t1,
                             "photons_per_s": fl["photons"] / fl["kernel_ms"] * 1e3,
                             "steps_per_photon": fl["totals"]["steps"] / max(1, fl["photons"])},
            "fluorescence_importance_sampled": {"photons": n_is, "device_ms": kms_is, "wall_s": t4 - t3,
                                                "FlR_total": is_tot, "FlR_total_sigma": is_sig,
                                                "FlR_total_every_voxel": float(fl["FlR"].sum())},
            "total_wall_s": t2 - t0}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--photons", type=float, default=1e8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, help="copy this many bytes device-to-device first")
    ap.add_argument("--scan", action="store_true", help="also run the 64-source batched scan")
    ap.add_argument("--hetero", action="store_true", help="also run the 32-medium heterogeneous batch")
    ap.add_argument("--serve", action="store_true", help="also run the served request stream")
    ap.add_argument("--sample2", type=float, default=0.0,
                    help="also run the reference's sample2.mci end to end, photon counts scaled by this factor")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    import cudamcfl_amd as M
    from cudamcfl_amd.media import inclusion_simulation

    N = int(args.photons)
    sim = inclusion_simulation(photons=N)
    ctx = M.Context(0)
    if args.copy_probe:
        copy_probe(args.copy_probe)
    info = ctx.info()
    for w in range(args.warmup):
        ctx.run_excitation(sim, seed=1000 + w)
    kms, walls, steps = [], [], 0
    t_all = time.perf_counter()
    for k in range(args.steps):
        t0 = time.perf_counter()
        r = ctx.run_excitation(sim, seed=k + 1)
        walls.append(time.perf_counter() - t0)
        kms.append(r.kernel_ms)
        print(f"step {k}: kernel {r.kernel_ms:.1f} ms, {N / r.kernel_ms * 1e3:.3e} photons/s", file=sys.stderr,
              flush=True)
        steps += r.totals["steps"]
        t = r.totals
        assert t["photons"] == N
        assert t["launched_w"] + t["roul_gain"] - t["roul_loss"] - t["zero_kill"] - t["aborted"] == \
            t["R"] + t["T"] + t["A"], "energy balance violated"
    t_all = time.perf_counter() - t_all
    dev_s = sum(kms) / 1e3
    value = N * args.steps / dev_s
    steps_per_s = steps / dev_s
    achieved = steps_per_s * F_STEP / 1e12
    peak = PEAK_SLOTS / 1e12
    traffic = measured_traffic(N)
    ms_step = sum(kms) / args.steps
    # the secondary roofline: HBM bytes per launch over the launch time, against 8 TB/s
    hbm = None
    if traffic:
        gbs = traffic / (ms_step * 1e-3) / 1e9
        hbm = {"achieved": gbs, "peak": 8000.0, "unit": "GB/s", "frac": gbs / 8000.0}
    out = {
        "metric": "photons/s at 128x128x64 voxels",
        "value": value,
        "unit": "photons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "wall_ms_per_step": 1e3 * sum(walls) / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain excitation: 128x128x64 uint8 labels (0.5 mm), 5 spheres + 2 cylinders, "
                               "8 materials, pencil beam, 1e8 photons, fhd + 128x128 R/T images",
                   "photons": N, "voxels": [128, 128, 64], "rng": "photon streams", "medium_seed": 2024,
                   "parallelism": "1 GPU"},
        "photon_steps_per_s": steps_per_s,
        "steps_per_photon": steps / (N * args.steps),
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "Tslot/s",
                     "frac": achieved / peak, "traffic": traffic, "hbm": hbm,
                     "note": f"{F_STEP} VALU issue slots per photon-step x photon-steps / kernel time; traffic = "
                             "HBM bytes per launch from FETCH_SIZE/WRITE_SIZE scaled on a copy (profiles/)"},
        "gpu": info,
    }
    if args.scan:
        out["batched_scan"] = bench_scan(ctx)
    if args.hetero:
        out["batched_hetero"] = bench_hetero(ctx)
    if args.serve:
        out["serving"] = bench_serve(ctx)
    if args.sample2 > 0:
        out["sample2_end_to_end"] = bench_sample2(ctx, args.sample2)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
