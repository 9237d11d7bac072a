"""The allocation-free excitation (mcfl_run_excitation_noalloc, every device
buffer a slice of one fixed scratch array per GPU) against
mcfl_run_excitation on the same inputs: integer tallies, totals and advanced
lane states bit-identical, on consecutive calls whose second input is
smaller than the first (the scratch keeps the previous call's contents), and
sizes beyond the scratch's limits refused before any launch."""
import numpy as np
import pytest

import cudamcfl_amd as M
from cudamcfl_amd.media import inclusion_simulation
from media import small_sim

pytestmark = pytest.mark.gpu


def _same(a, b):
    for n in ("Rd", "Tt", "fhd", "A_xyz", "time", "Rd2", "Tt2"):
        x, y = getattr(a, n, None), getattr(b, n, None)
        if x is None and y is None:
            continue
        assert np.array_equal(x, y), n
    assert a.totals == b.totals and a.totals["photons"] > 0


def test_noalloc_equals_the_allocating_entry_point(ctx):
    big = inclusion_simulation(photons=300_000)  # 2^20 voxels: the scratch's limit for fhd
    print("case: noalloc, 128x128x64, 3e5 photons, fhd + squared weights", flush=True)
    a = ctx.run_excitation(big, seed=3, sq=True)
    b = ctx.run_excitation(big, seed=3, sq=True, noalloc=True)
    _same(a, b)
    assert b.fhd.sum() > 0
    small = small_sim(photons=20000)  # 20x20x10, temporal detectors, fhd, A_xyz
    print("case: noalloc, 20x20x10, 2e4 photons, fhd + A_xyz + squared weights + time", flush=True)
    a = ctx.run_excitation(small, seed=5, axyz=True, sq=True)
    b = ctx.run_excitation(small, seed=5, axyz=True, sq=True, noalloc=True)
    _same(a, b)
    assert b.time.sum() > 0 and b.A_xyz.sum() > 0
    print("case: noalloc, 20x20x10, 512 reference lanes", flush=True)
    sl = small_sim(photons=3001, temporal=False)
    x, am = M.init_rng(512, seed=77)
    xa, xb = x.copy(), x.copy()
    a = ctx.run_excitation(sl, lanes=(xa, am))
    b = ctx.run_excitation(sl, lanes=(xb, am), noalloc=True)
    _same(a, b)
    assert np.array_equal(xa, xb) and not np.array_equal(xa, x)


def test_noalloc_refuses_sizes_beyond_the_scratch(ctx):
    big = inclusion_simulation(photons=1000)
    with pytest.raises(M.McflError, match="capacity"):  # fhd and A_xyz copies of 2^20 voxels: twice the region
        ctx.run_excitation(big, seed=1, axyz=True, noalloc=True)
    big.s.number_of_photons = 2 ** 27 + 1  # a split run
    with pytest.raises(M.McflError, match="capacity"):
        ctx.run_excitation(big, seed=1, noalloc=True)
    # the refusals leave the scratch usable
    small = small_sim(photons=5000)
    _same(ctx.run_excitation(small, seed=9), ctx.run_excitation(small, seed=9, noalloc=True))
