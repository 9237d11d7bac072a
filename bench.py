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


def cpu_baseline(target_s=12.0):
    """The CPU oracle on host cores, photon streams, same medium; a bounded
    sample sized to about target_s seconds."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import cudamcfl_amd as M
    import _oracle as O
    from cudamcfl_amd.media import cfg2_simulation

    threads = min(16, os.cpu_count() or 1)
    sim = cfg2_simulation(photons=10_000)
    o = O.FromProduct(sim)
    a_tab = M.multipliers(32768)
    key = M.stream_key(99, 0)
    t0 = time.perf_counter()
    r = O.run_photons(o.o, a_tab, key, 0, 4000, nthreads=threads)
    dt = time.perf_counter() - t0
    n = int(max(4000, min(5_000_000, 4000 * target_s / max(dt, 1e-3))))
    t0 = time.perf_counter()
    r = O.run_photons(o.o, a_tab, key, 0, n, nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "photons/s", "cores": threads, "kind": "port",
            "sample": f"{n} photons of the same 128x128x64 medium (photon streams), {dt:.1f} s",
            "photon_steps_per_s": r.tot["steps"] / dt}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--photons", type=float, default=1e8)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    import cudamcfl_amd as M
    from cudamcfl_amd.media import cfg2_simulation

    N = int(args.photons)
    sim = cfg2_simulation(photons=N)
    ctx = M.Context(0)
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
    out = {
        "metric": "photons/s at 128x128x64 voxels",
        "value": value,
        "unit": "photons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sum(kms) / args.steps,
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
                     "frac": achieved / peak, "traffic": None,
                     "note": f"{F_STEP} VALU issue slots per photon-step x photon-steps / kernel time"},
        "gpu": info,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
