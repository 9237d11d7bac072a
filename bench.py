"""Benchmark of the voxelized fluorescence Monte Carlo path on one MI355X.

Headline (`value`): the plain path.  The 128x128x64-voxel seeded medium (0.5
mm voxels, 5 spheres + 2 cylinders, 8 materials), pencil beam at the centre,
1e8 photons per step, photon hitting density (3D) and 128x128
reflectance/transmittance images tallied.  One step = one whole simulation
= one persistent kernel launch.  `value` is photons per second of device
time with the medium resident in HBM (HIP events on the launch stream); the
wall time including host<->device copies is reported beside it.

The other half of the metric, in the same line: the 64-source scan batched
in one launch against 64 plain launches and against the drop-in workflow of
one `mcfl` process per .mci file, and a served Poisson request stream
(p50/p99 latency) against a one-launch-per-request loop; each form with its
photons/s and photon-steps/s.  `cpu_baseline`: the oracle built for the host
(-O3 -march=native) on the host's cores.

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

# VALU issue slots per photon-step of the reference loop (CUDAMCFLtransport.cu:
# 286-447, counted with transcendentals and 64-bit multiply-adds at 4 slots).
# Of these, the time-of-flight update (6) only feeds the temporal tally and
# the fhd index and guard (11) only the photon hitting density: a launch
# without them needs fewer.
F_STEP_REF, F_TOF, F_FHD = 150, 6, 11


def f_step(temporal, fhd):
    return F_STEP_REF - (0 if temporal else F_TOF) - (0 if fhd else F_FHD)


# MI355X VALU issue ceiling in lane slots/s: the published 157.3 TFLOP/s FP32
# vector figure / 2 flops per FMA = 256 CU x 4 SIMD x 32 lanes x 2.4 GHz, one
# wave64 instruction per 2 cycles per SIMD.  Streams of independent FP32 or
# integer instructions with vector-register operands issue at 2.4-2.8 cycles
# per wave64 instruction on the GPU (probe), below it; an operand in a scalar
# register costs 4.4 cycles when every instruction has one.
PEAK_SLOTS = 256 * 4 * 32 * 2.4e9
CPU_THREADS_MAX = 16  # the host CPU share of one GPU on the measuring boxes


def _host_threads():
    """Threads for the CPU baseline: the cores this process may run on, at
    most the host's CPU share of one GPU."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return min(avail, CPU_THREADS_MAX), avail


def _native_oracle():
    """The oracle compiled for this host (-O3 -march=native, FP contraction
    off, no fast-math) into a temporary directory; None if gcc fails."""
    import subprocess
    import tempfile
    d = tempfile.mkdtemp(prefix="mcfl_oracle_native_")
    so = os.path.join(d, "liboracle.so")
    cmd = ["gcc", "-O3", "-march=native", "-ffp-contract=off", "-fno-math-errno", "-fopenmp", "-fPIC", "-std=gnu11",
           "-shared", "-o", so, os.path.join(ROOT, "oracle", "mcfl_oracle.c"), "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return (so, " ".join(cmd[:4])) if r.returncode == 0 else (None, r.stderr[-300:])


def _timed_sample(O, sim_o, a_tab, key, threads, target_s):
    """The oracle on growing photon counts until a run takes ~target_s."""
    n = 2000
    for _ in range(4):  # calibrate on growing runs (thread start-up dominates tiny ones)
        t0 = time.perf_counter()
        O.run_photons(sim_o, a_tab, key, 0, n, nthreads=threads)
        dt = time.perf_counter() - t0
        if dt > 0.5:
            break
        n *= 8
    n = int(max(2000, min(40_000_000, n * target_s / max(dt, 1e-3))))
    t0 = time.perf_counter()
    r = O.run_photons(sim_o, a_tab, key, 0, n, nthreads=threads)
    return n, time.perf_counter() - t0, r


def cpu_baseline(target_s=20.0):
    """The CPU oracle (photon streams) on the host's cores, built for this host
    at -O3 -march=native: the bench medium (bounded sample of ~target_s s) and
    the small 80x80x20 sphere problem, each also on one thread.  The native
    build's tallies are checked equal to the shipped build's on a small run."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import cudamcfl_amd as M
    import _oracle as O
    from cudamcfl_amd.media import inclusion_simulation, sphere_simulation

    threads, avail = _host_threads()
    a_tab = M.multipliers(32768)
    sim2 = inclusion_simulation(photons=10_000)
    o2 = O.FromProduct(sim2)
    so, build = _native_oracle()
    same_as_shipped = None
    if so:
        ref = O.run_photons(o2.o, a_tab, 12345, 0, 4000, nthreads=threads)  # the shipped -O2 build
        O._lib, O.LIB_PATH = None, so
        nat = O.run_photons(o2.o, a_tab, 12345, 0, 4000, nthreads=threads)
        same_as_shipped = bool(np.array_equal(ref.Rd, nat.Rd) and np.array_equal(ref.fhd, nat.fhd) and
                               ref.tot == nat.tot)
    else:
        build = "shipped -O2 -march=x86-64-v3 build (native build failed: %s)" % build
    key = M.stream_key(99, 0)
    n, dt, r = _timed_sample(O, o2.o, a_tab, key, threads, target_s)
    n1, dt1, r1 = _timed_sample(O, o2.o, a_tab, key, 1, 3.0)
    sim1 = sphere_simulation(photons=10_000)
    o1 = O.FromProduct(sim1)
    m, dtm, rm = _timed_sample(O, o1.o, a_tab, key, threads, target_s / 2)
    m1, dtm1, rm1 = _timed_sample(O, o1.o, a_tab, key, 1, 3.0)
    return {"value": n / dt, "unit": "photons/s", "cores": threads, "kind": "port",
            "sample": f"{n} photons of the bench medium (128x128x64, photon streams, fhd on), {dt:.1f} s",
            "photon_steps_per_s": r.tot["steps"] / dt, "steps_per_photon": r.tot["steps"] / n,
            "build": build, "native_build_equals_shipped": same_as_shipped,
            "host_cpus_available": avail,
            "single_thread": {"photons_per_s": n1 / dt1, "photon_steps_per_s": r1.tot["steps"] / dt1,
                              "sample": f"{n1} photons, {dt1:.1f} s"},
            "small_sphere": {"workload": "80x80x20 voxels of 1 mm, sphere r 0.5 cm, pencil beam (the small "
                                         "plain configuration)",
                             "photons_per_s": m / dtm, "photon_steps_per_s": rm.tot["steps"] / dtm,
                             "sample": f"{m} photons, {dtm:.1f} s, {threads} threads",
                             "single_thread_photons_per_s": m1 / dtm1}}


def measured_counters(photons):
    """Per-launch counters of this workload from the committed profile: HBM
    bytes (FETCH_SIZE and WRITE_SIZE, each scaled on a plain copy) and VALU
    wave-instructions issued (SQ_INSTS_VALU); None when the profile covers
    another photon count."""
    path = os.path.join(ROOT, "profiles", "mcfl_transport_kernel_128x128x64_1e8_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    return t if t.get("photons") == photons else None


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


def rse_stats(jobs, frac):
    """Per-pixel relative standard error of the reflectance images from the
    sum and sum-of-squares tallies, RSE = sqrt(Q - S^2/N) / S, over the
    pixels above `frac` of each image's peak."""
    vals = []
    for j in jobs:
        S = j.Rd.astype(np.float64)
        Q = j.Rd2.astype(np.float64) * 2.0 ** 32
        keep = S > frac * S.max()
        vals.append(np.sqrt(np.maximum(Q[keep] - S[keep] ** 2 / j.photons, 0.0)) / S[keep])
    v = np.concatenate(vals)
    return {"above_frac_of_peak": frac, "pixels": int(v.size), "median": float(np.median(v)),
            "p90": float(np.percentile(v, 90)), "max": float(v.max()), "frac_le_1pct": float(np.mean(v <= 0.01))}


def bench_scan(ctx, lanes=False):
    """64-source scan on the 128x128x64 medium: one graph-captured batched
    launch (autotuned) vs 64 plain launches with the same budgets and seeds
    (bit-identical images, hence identical per-pixel RSE)."""
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
    steps_b = sum(j.totals["steps"] for j in jobs)
    print(f"scan batched: {wall_b:.2f} s wall, {kms:.0f} ms device", file=sys.stderr, flush=True)
    plain_wall, plain_ms, steps_p, same = 0.0, 0.0, 0, True
    for i, (s, n) in enumerate(zip(srcs, budgets)):
        sim.s.number_of_photons = int(n)
        sim.s.xi, sim.s.yi, sim.s.zi = s
        t0 = time.perf_counter()
        r = ctx.run_excitation(sim, seed=700 + i, sq=True)
        plain_wall += time.perf_counter() - t0
        plain_ms += r.kernel_ms
        steps_p += r.totals["steps"]
        same &= bool(np.array_equal(r.Rd, jobs[i].Rd) and np.array_equal(r.Rd2, jobs[i].Rd2))
    print(f"scan plain: {plain_wall:.2f} s wall, {plain_ms:.0f} ms device, identical={same}", file=sys.stderr,
          flush=True)
    out = {"workload": "8x8 pencil scan, 8-voxel pitch, budgets log-uniform [1e6,5e7] (seed 7), 128x128x64 medium, "
                       "R/T images and per-pixel sums of squared weights",
           "photons": total,
           "batched": {"photons_per_s": total / wall_b, "device_ms": kms, "wall_s": wall_b,
                       "photon_steps_per_s": steps_b / (kms * 1e-3), "photon_steps": steps_b},
           "plain_launches": {"photons_per_s": total / plain_wall, "device_ms": plain_ms, "wall_s": plain_wall,
                              "photon_steps_per_s": steps_p / (plain_ms * 1e-3), "photon_steps": steps_p},
           "speedup_wall": plain_wall / wall_b, "speedup_device": plain_ms / kms,
           "images_identical": same, "rse": [rse_stats(jobs, 0.1), rse_stats(jobs, 0.01)], "autotune": tune}
    if lanes:  # the reference's launch geometry: 32768 MWC lanes per simulation (DoOneSimulation)
        lanes_wall, lanes_ms = 0.0, 0.0
        for i, (s, n) in enumerate(zip(srcs, budgets)):
            sim.s.number_of_photons = int(n)
            sim.s.xi, sim.s.yi, sim.s.zi = s
            x, a = M.init_rng(M.NUM_THREADS, seed=900 + i)
            t0 = time.perf_counter()
            r = ctx.run_excitation(sim, lanes=(x, a), sq=True)
            lanes_wall += time.perf_counter() - t0
            lanes_ms += r.kernel_ms
        out["plain_reference_lanes"] = {"lanes": M.NUM_THREADS, "photons_per_s": total / lanes_wall,
                                        "device_ms": lanes_ms, "wall_s": lanes_wall}
    b.close()
    return out


def bench_dropin(n_files=64):
    """The drop-in workflow of the 64-source scan: one `mcfl <file>.mci -S7`
    process per source (the reference's one-process-per-file scans) against
    one `mcfl f1.mci ... f64.mci -S7` process (the files batched), wall time
    of the processes included; the image files are compared byte for byte."""
    import subprocess
    import tempfile
    from cudamcfl_amd.media import inclusion_medium, scan_budgets, scan_sources, write_bulk, write_mci
    exe = os.path.join(ROOT, "cudamcfl_amd", "_build", "mcfl")
    lab, mats = inclusion_medium(2024)
    nz, ny, nx = lab.shape
    srcs, budgets = scan_sources()[:n_files], scan_budgets()[:n_files]
    base = tempfile.mkdtemp(prefix="mcfl_dropin_")
    for d in ("one_each", "batched"):
        os.makedirs(os.path.join(base, d))
        write_bulk(os.path.join(base, d, "medium.bulk"), lab)
        for i, (src, n) in enumerate(zip(srcs, budgets)):
            write_mci(os.path.join(base, d, f"s{i}.mci"), f"s{i}", "medium.bulk", mats, esp=nz / 20, grid_size=20,
                      photons=int(n), do_fl=0, det=(0.05, 0.05, nx, ny, 0.0, 0.0), source=src, dims=(nx, ny, nz))
    names = [f"s{i}.mci" for i in range(len(srcs))]
    t0 = time.perf_counter()
    for f in names:
        r = subprocess.run([exe, f, "-S7"], cwd=os.path.join(base, "one_each"), capture_output=True, text=True)
        assert r.returncode == 0, r.stdout[-2000:]
    wall_each = time.perf_counter() - t0
    print(f"drop-in: {len(names)} processes, {wall_each:.1f} s", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    r = subprocess.run([exe] + names + ["-S7"], cwd=os.path.join(base, "batched"), capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:]
    wall_one = time.perf_counter() - t0
    print(f"drop-in: one process, {wall_one:.1f} s", file=sys.stderr, flush=True)
    same = all(open(os.path.join(base, "one_each", f"s{i}{suf}"), "rb").read() ==
               open(os.path.join(base, "batched", f"s{i}{suf}"), "rb").read()
               for i in range(len(names)) for suf in ("_Refl.dat", "_Trans.dat"))
    total = int(sum(budgets))
    return {"workload": f"{len(names)} .mci files of the scan (128x128x64 bulk file, budgets log-uniform [1e6,5e7])",
            "photons": total, "one_process_per_file": {"wall_s": wall_each, "photons_per_s": total / wall_each},
            "one_process_all_files": {"wall_s": wall_one, "photons_per_s": total / wall_one},
            "speedup_wall": wall_each / wall_one, "image_files_identical": bool(same)}


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
    the server steps whenever it has work.  Returns latencies, span, photon
    steps of the requests, server counters."""
    import cudamcfl_amd as M
    srv = M.Server(ctx, mem_budget=48 << 30, epoch_photons=epoch)
    mids = [srv.add_medium(m) for m in media]
    t0 = time.perf_counter()
    arr, lat, i, steps = {}, {}, 0, 0
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
            out = srv.collect(rid, arr[rid][1])
            steps += out["totals"]["steps"] + out["fl_totals"]["steps"]
    span = time.perf_counter() - t0
    cnt = srv.counters()
    srv.close()
    return np.array(sorted(lat.values())), span, steps, cnt


def _plain(ctx, media, reqs):
    """One launch per request (+ one emission launch for a fluorescence
    request) in arrival order on the same device.  Returns latencies, span,
    photon steps and device time."""
    import cudamcfl_amd as M
    t0 = time.perf_counter()
    lat, steps, dev_ms = [], 0, 0.0
    for k, (ta, m, src, n, seed, fl) in enumerate(reqs):
        now = time.perf_counter() - t0
        if now < ta:
            time.sleep(ta - now)
        s = media[m]
        s.s.number_of_photons = n
        s.s.xi, s.s.yi, s.s.zi = src
        s.s.src_kind = M.SOURCE_PENCIL
        s.s.fhd_activated = 1 if fl else 0
        r = ctx.run_excitation(s, seed=seed)
        steps += r.totals["steps"]
        dev_ms += r.kernel_ms
        if fl:
            Fx = r.fhd.astype(np.float64) / (4294967295.0 * r.totals["photons"])
            job = M.BatchJob(0, n, seed, src_kind=M.SOURCE_EMISSION, Fx=Fx)
            bb = M.Batch(ctx, [s], [job])
            dev_ms += bb.run()
            bb.close()
            steps += job.totals["steps"]
        lat.append(time.perf_counter() - t0 - ta)
    return np.array(sorted(lat)), time.perf_counter() - t0, steps, dev_ms


def bench_serve(ctx, rate=12.0, duration=8.0, epoch=250_000):
    """A Poisson request stream (the served configuration's generator: media
    up to 128x128x64, scan sources, half with fluorescence, budgets
    log-uniform [1e5,1e8]): continuous batching against a plain
    one-launch-per-request FIFO loop on the same arrivals and device."""
    from cudamcfl_amd.media import request_stream

    media, reqs = request_stream(rate=rate, duration=duration)
    work = sum(n * (2 if fl else 1) for (_, _, _, n, _, fl) in reqs)
    served, span_s, steps_s, cnt = _served(ctx, media, reqs, epoch)
    print(f"served at {rate}/s: {len(served)} requests, span {span_s:.1f} s", file=sys.stderr, flush=True)
    plain, span_p, steps_p, dev_p = _plain(ctx, media, reqs)
    print(f"plain at {rate}/s: span {span_p:.1f} s", file=sys.stderr, flush=True)

    def lat(l):
        return {"p50_s": float(np.percentile(l, 50)), "p99_s": float(np.percentile(l, 99)),
                "mean_s": float(np.mean(l))}

    return {"workload": f"Poisson arrivals {rate}/s for {duration} s ({len(reqs)} requests), 8 media up to 128x128x64, "
                        f"scan sources, 50% with fluorescence, budgets log-uniform [1e5,1e8] (seed 5)",
            "photons": int(work), "epoch_photons": epoch,
            "served": dict(lat(served), photons_per_s=work / span_s, span_s=span_s, epochs=cnt["epochs"],
                           device_photons_per_s=cnt["photons"] / (cnt["kernel_ms"] * 1e-3),
                           device_photon_steps_per_s=steps_s / (cnt["kernel_ms"] * 1e-3),
                           kernel_s=cnt["kernel_ms"] / 1e3, host_s=cnt["host_ms"] / 1e3),
            "plain_loop": dict(lat(plain), photons_per_s=work / span_p, span_s=span_p,
                               device_photons_per_s=work / (dev_p * 1e-3),
                               device_photon_steps_per_s=steps_p / (dev_p * 1e-3), kernel_s=dev_p / 1e3),
            "p50_ratio_plain_over_served": float(np.percentile(plain, 50) / np.percentile(served, 50)),
            "p99_ratio_plain_over_served": float(np.percentile(plain, 99) / np.percentile(served, 99))}


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
    ex_sps = ex.totals["steps"] / (ex.kernel_ms * 1e-3)
    fl_sps = fl["totals"]["steps"] / (fl["kernel_ms"] * 1e-3)
    F_ex, F_fl = f_step(temporal=True, fhd=True), f_step(temporal=False, fhd=False)
    return {"workload": f"sample2.mci ({s.nx}x{s.ny}x{s.nz} voxels from 3dbulk-sample, temporal detectors), "
                        f"photon counts x{frac:g}",
            "excitation": {"photons": n_ex, "device_ms": ex.kernel_ms, "wall_s": t1 - t0,
                           "photons_per_s": n_ex / ex.kernel_ms * 1e3,
                           "steps_per_photon": ex.totals["steps"] / n_ex, "photon_steps_per_s": ex_sps,
                           "F_step": F_ex, "valu_frac": ex_sps * F_ex / PEAK_SLOTS},
            "fluorescence": {"voxels": fl["voxels"], "photons_per_voxel": n_v, "photons": fl["photons"],
                             "device_ms": fl["kernel_ms"], "wall_s": t2 - t1,
                             "photons_per_s": fl["photons"] / fl["kernel_ms"] * 1e3,
                             "steps_per_photon": fl["totals"]["steps"] / max(1, fl["photons"]),
                             "photon_steps_per_s": fl_sps, "F_step": F_fl, "valu_frac": fl_sps * F_fl / PEAK_SLOTS},
            "fluorescence_importance_sampled": {"photons": n_is, "device_ms": kms_is, "wall_s": t4 - t3,
                                                "FlR_total": is_tot, "FlR_total_sigma": is_sig,
                                                "FlR_total_every_voxel": float(fl["FlR"].sum())},
            "total_wall_s": t2 - t0}


def bench_small_plain(ctx, photons=1_000_000):
    """The small plain configuration on the GPU: the reference-shaped
    80x80x20 sphere problem, 1e6 photons, fhd and 40x40 images."""
    from cudamcfl_amd.media import sphere_simulation
    sim = sphere_simulation(photons=photons)
    ctx.run_excitation(sim, seed=1)
    r = ctx.run_excitation(sim, seed=2)
    return {"workload": "80x80x20 voxels of 1 mm, sphere r 0.5 cm at 1 cm, pencil beam, 1e6 photons, fhd + 40x40 "
                        "images", "photons": photons, "device_ms": r.kernel_ms,
            "photons_per_s": photons / (r.kernel_ms * 1e-3),
            "photon_steps_per_s": r.totals["steps"] / (r.kernel_ms * 1e-3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--photons", type=float, default=1e8)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-scan", action="store_true", help="skip the 64-source batched scan")
    ap.add_argument("--no-serve", action="store_true", help="skip the served request stream")
    ap.add_argument("--no-dropin", action="store_true", help="skip the one-process-per-file scan")
    ap.add_argument("--plain-only", action="store_true",
                    help="the plain path alone (with --hetero / --sample2 if given); for profiling")
    ap.add_argument("--scan-lanes", action="store_true", help="the scan also in the reference's 32768-lane geometry")
    ap.add_argument("--copy-probe", type=int, default=0, help="copy this many bytes device-to-device first")
    ap.add_argument("--hetero", action="store_true", help="also run the 32-medium heterogeneous batch")
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
    dev_s = sum(kms) / 1e3
    value = N * args.steps / dev_s
    steps_per_s = steps / dev_s
    F = f_step(temporal=False, fhd=True)
    achieved = steps_per_s * F / 1e12
    peak = PEAK_SLOTS / 1e12
    prof = measured_counters(N)
    traffic = prof["bytes_per_launch"] if prof else None
    ms_step = sum(kms) / args.steps
    # the secondary roofline: HBM bytes per launch over the launch time, against 8 TB/s
    hbm = issued = None
    if traffic:
        gbs = traffic / (ms_step * 1e-3) / 1e9
        hbm = {"achieved": gbs, "peak": 8000.0, "unit": "GB/s", "frac": gbs / 8000.0}
    if prof and prof.get("valu_insts"):
        # lane slots the kernel issues (VALU wave-instructions x 64) over this launch's time: at most the ceiling
        ts = prof["valu_insts"] * 64 / (ms_step * 1e-3) / 1e12
        issued = {"achieved": ts, "frac": ts / peak, "slots_per_step": prof["valu_insts"] * 64 / prof["steps"],
                  "source": "SQ_INSTS_VALU of the committed counter profile, over this run's launch time"}
    out = {
        "metric": "photons/s at 128x128x64 voxels; p50/p99 sim latency, batched vs plain",
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
                     "frac": achieved / peak, "traffic": traffic, "hbm": hbm, "issued": issued,
                     "F_step": F, "F_step_reference": F_STEP_REF,
                     "ceiling": "VALU issue, 256 CU x 4 SIMD x 32 lanes x 2.4 GHz = 78.6 Tslot/s (the published "
                                "157.3 TFLOP/s FP32 vector figure / 2 flops per FMA): one wave64 instruction per 2 "
                                "cycles per SIMD",
                     "note": f"achieved = {F} algorithmic VALU slots per photon-step (the reference loop's "
                             f"{F_STEP_REF} less the time-of-flight update this launch does not need: no temporal "
                             "detectors) x photon-steps / kernel time; issued = the slots the kernel actually issues "
                             "(divergent branches, addresses, 64-bit tallies); traffic = HBM bytes per launch from "
                             "FETCH_SIZE/WRITE_SIZE scaled on a copy (profiles/)"},
        "gpu": info,
    }
    if not args.plain_only:
        out["small_plain"] = bench_small_plain(ctx)
        if not args.no_scan:
            out["batched_scan"] = bench_scan(ctx, lanes=args.scan_lanes)
        if not args.no_dropin:
            out["dropin_scan"] = bench_dropin()
        if not args.no_serve:
            out["serving"] = bench_serve(ctx)
    if args.hetero:
        out["batched_hetero"] = bench_hetero(ctx)
    if args.sample2 > 0:
        out["sample2_end_to_end"] = bench_sample2(ctx, args.sample2)
    if not (args.plain_only or args.no_cpu):
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
