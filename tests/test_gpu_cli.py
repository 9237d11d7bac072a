"""The `mcfl` command on the reference's own method-1 input (sample.mci, a
reduced photon budget): it writes the reference's files, and its image files
equal those of the same run made through the C ABI and written by the same
writer."""
import ctypes as C
import pathlib
import subprocess

import pytest

import cudamcfl_amd as M

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).parent / "golden" / "ref_inputs"
EXE = pathlib.Path(M.__file__).parent / "_build" / "mcfl"


def test_cli_sample_mci(tmp_path, ctx):
    text = (GOLDEN / "sample.mci").read_text()
    assert "500000000" in text
    (tmp_path / "sample.mci").write_text(text.replace("500000000", "60000", 1))
    print("case: mcfl CLI, sample.mci with 60000 photons, -S7", flush=True)
    r = subprocess.run([str(EXE), "sample.mci", "-S7", "-F3"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "All done, no errors" in r.stdout
    out = "sample1-out"
    for suffix in ("_param.dat", "_Refl.dat", "_Trans.dat", "_time.dat", "_PHD-Ascii.dat", "_FlR.dat", "_FlT.dat"):
        assert (tmp_path / (out + suffix)).stat().st_size > 0, suffix

    api = tmp_path / "api"
    api.mkdir()
    sim = M.Simulation.from_mci(str(tmp_path / "sample.mci"))
    sim.s.outp_filename = str(api / out).encode()
    res = ctx.run_excitation(sim, seed=7)
    assert res.totals["photons"] == 60000
    assert M.lib().mcfl_write_simulation_results(C.byref(sim.s), C.byref(res.t), 0.0) == 0
    for suffix in ("_Refl.dat", "_Trans.dat", "_time.dat"):
        assert (tmp_path / (out + suffix)).read_bytes() == (api / (out + suffix)).read_bytes(), suffix


def test_cli_several_files_equal_single_runs(tmp_path):
    """`mcfl a.mci b.mci c.mci -S5` (two files on one medium, one on another,
    temporal detectors, fhd and fluorescence) writes, for every file, the same
    image, time and PHD files as `mcfl <file> -S5` run alone."""
    from media import MATS3, small_labels
    from cudamcfl_amd.media import write_bulk, write_mci
    lab_a, lab_b = small_labels(20, 20, 10, 0), small_labels(16, 16, 8, 1)
    mats = [[m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7]] for m in MATS3]
    tmp = (0.0, 0.0, 2, 1, 0.4, 0.4, 0.3, 0, 50, 1000000)
    for d in ("many", "one"):
        (tmp_path / d).mkdir()
        write_bulk(tmp_path / d / "a.bulk", lab_a)
        write_bulk(tmp_path / d / "b.bulk", lab_b)
        for name, bulk, lab, src in (("f1", "a.bulk", lab_a, (0.0, 0.0, 0.0)), ("f2", "a.bulk", lab_a, (0.3, -0.2, 0.0)),
                                     ("f3", "b.bulk", lab_b, (0.1, 0.1, 0.0))):
            nz, ny, nx = lab.shape
            write_mci(tmp_path / d / f"{name}.mci", name, bulk, mats, esp=nz / 10, grid_size=10, photons=40000,
                      photons_per_voxel=33000, do_fl=1, det=(0.1, 0.1, nx, ny, 0.0, 0.0), source=src, n_below=1.33,
                      dims=(nx, ny, nz), temporal=tmp)
    print("case: mcfl CLI, 3 files in one batch vs alone, -S5 -F9", flush=True)
    r = subprocess.run([str(EXE), "f1.mci", "f2.mci", "f3.mci", "-S5", "-F9"], cwd=tmp_path / "many",
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "in one batch" in r.stdout
    for name in ("f1", "f2", "f3"):
        r1 = subprocess.run([str(EXE), f"{name}.mci", "-S5", "-F9"], cwd=tmp_path / "one", capture_output=True,
                            text=True, timeout=300)
        assert r1.returncode == 0, r1.stdout[-3000:] + r1.stderr[-3000:]
        for suffix in ("_Refl.dat", "_Trans.dat", "_time.dat", "_PHD-Ascii.dat", "_FlR.dat", "_FlT.dat"):
            a = (tmp_path / "many" / (name + suffix)).read_bytes()
            b = (tmp_path / "one" / (name + suffix)).read_bytes()
            assert a == b, (name, suffix)


def test_cli_importance_sampled_emission(tmp_path):
    """`mcfl f.mci -I<N>` estimates the same fluorescence images as the
    every-voxel loop: the FlR totals agree within Monte Carlo noise."""
    import numpy as np
    from media import MATS3, small_labels
    from cudamcfl_amd.media import write_bulk, write_mci
    from cudamcfl_amd.files import read_image
    lab = small_labels(20, 20, 10, 2)
    nz, ny, nx = lab.shape
    for d in ("loop", "is"):
        (tmp_path / d).mkdir()
        write_bulk(tmp_path / d / "a.bulk", lab)
        write_mci(tmp_path / d / "f.mci", "f", "a.bulk", MATS3, esp=nz / 10, grid_size=10, photons=200000,
                  photons_per_voxel=33000, do_fl=1, det=(0.1, 0.1, nx, ny, 0.0, 0.0), n_below=1.33, dims=(nx, ny, nz))
    print("case: mcfl CLI, every-voxel emission vs -I4000000", flush=True)
    r = subprocess.run([str(EXE), "f.mci", "-S5", "-F9"], cwd=tmp_path / "loop", capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    r = subprocess.run([str(EXE), "f.mci", "-S5", "-F9", "-I4000000"], cwd=tmp_path / "is", capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "importance-sampled emission" in r.stdout
    a = read_image(str(tmp_path / "loop" / "f_FlR.dat"))
    b = read_image(str(tmp_path / "is" / "f_FlR.dat"))
    assert a.shape == b.shape and a.sum() > 0 and b.sum() > 0
    assert abs(a.sum() - b.sum()) < 0.03 * a.sum(), (a.sum(), b.sum())
    # the excitation outputs do not depend on how the emission stage is run
    for suffix in ("_Refl.dat", "_Trans.dat", "_PHD-Ascii.dat"):
        assert (tmp_path / "loop" / ("f" + suffix)).read_bytes() == (tmp_path / "is" / ("f" + suffix)).read_bytes()


def test_cli_lanes_both_stages(tmp_path, ctx, monkeypatch):
    """`mcfl f.mci -L -S5 -F9` runs both stages on the reference's 32768
    lane streams (excitation seeded with 5, emission reseeded with 9): its
    fluorescence files equal those of mcfl_DoOneSimulation followed by
    mcfl_run_fluorescence_lanes through the C ABI, written by the same writer."""
    import numpy as np
    from media import MATS3, small_labels
    from cudamcfl_amd.media import write_bulk, write_mci
    lab = small_labels(10, 10, 5, 3)
    nz, ny, nx = lab.shape
    write_bulk(tmp_path / "a.bulk", lab)
    write_mci(tmp_path / "f.mci", "f", "a.bulk", MATS3, esp=nz / 10, grid_size=10, photons=40000,
              photons_per_voxel=33000, do_fl=1, det=(0.1, 0.1, nx, ny, 0.0, 0.0), n_below=1.33, dims=(nx, ny, nz))
    print("case: mcfl CLI -L, both stages on 32768 reference lanes", flush=True)
    r = subprocess.run([str(EXE), "f.mci", "-L", "-S5", "-F9"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "All done, no errors" in r.stdout

    api = tmp_path / "api"
    api.mkdir()
    monkeypatch.chdir(tmp_path)  # the bulk file is named relative to the working directory
    sim = M.Simulation.from_mci("f.mci")
    sim.s.outp_filename = str(api / "f").encode()
    x, a = M.init_rng(M.NUM_THREADS, seed=5)
    Fx = np.zeros(nx * ny * nz)
    done = M.lib().mcfl_DoOneSimulation(ctx.h, C.byref(sim.s), x.ctypes.data_as(M.U64P), a.ctypes.data_as(M.U32P),
                                        Fx.ctypes.data_as(M.F64P))
    assert done == 40000
    x2, a2 = M.init_rng(M.NUM_THREADS, seed=9)
    fl = ctx.run_fluorescence_lanes(sim, Fx, 33000, lanes=(x2, a2))
    assert fl["raw_R"].sum() > 0
    M.lib().mcfl_write_fl(C.byref(sim.s), fl["FlR"].ctypes.data_as(M.F64P), fl["FlT"].ctypes.data_as(M.F64P))
    for suffix in ("_FlR.dat", "_FlT.dat"):
        assert (tmp_path / ("f" + suffix)).read_bytes() == (api / ("f" + suffix)).read_bytes(), suffix


def test_cli_lanes_on_the_layered_sample(tmp_path):
    """`mcfl sample.mci -L -S7` (method 1, the reference's 32768 lanes): the
    run finishes every photon and writes the reference's files; its image
    files equal those of mcfl_DoOneSimulation through the C ABI with the same
    seed (which writes them itself)."""
    text = (GOLDEN / "sample.mci").read_text()
    (tmp_path / "cli").mkdir()
    (tmp_path / "cli" / "sample.mci").write_text(text.replace("500000000", "60000", 1))
    print("case: mcfl CLI -L, sample.mci with 60000 photons", flush=True)
    r = subprocess.run([str(EXE), "sample.mci", "-L", "-S7"], cwd=tmp_path / "cli", capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "All done, no errors" in r.stdout
    api = tmp_path / "api"
    api.mkdir()
    sim = M.Simulation.from_mci(str(tmp_path / "cli" / "sample.mci"))
    sim.s.outp_filename = str(api / "sample1-out").encode()
    x, a = M.init_rng(M.NUM_THREADS, seed=7)
    ctx = M.Context(0)
    done = M.lib().mcfl_DoOneSimulation(ctx.h, C.byref(sim.s), x.ctypes.data_as(M.U64P), a.ctypes.data_as(M.U32P),
                                        None)
    ctx.close()
    assert done == 60000
    for suffix in ("_Refl.dat", "_Trans.dat"):
        assert (tmp_path / "cli" / ("sample1-out" + suffix)).read_bytes() == (api / ("sample1-out" + suffix)).read_bytes()


def test_cli_layered_files_in_one_batch(tmp_path):
    """`mcfl f1.mci f2.mci f3.mci -S5` on the reference's method-1 sample.mci
    with three source positions (photon budget reduced) runs the three
    excitations in one batched launch and writes, for every file, the image,
    time and PHD files of the same file run alone."""
    text = (GOLDEN / "sample.mci").read_text()
    lines = text.splitlines(keepends=True)
    src = next(i for i, l in enumerate(lines) if "ource" in l and not l.lstrip().startswith("#"))
    assert lines[src].startswith("-4 0 0"), lines[src]
    for d in ("many", "one"):
        (tmp_path / d).mkdir()
        for k, pos in enumerate(("-4 0 0", "-3.9 0.1 0", "-4.1 -0.05 0")):
            body = lines[:]
            body[src] = pos + " # Source position\n"
            t = "".join(body).replace("500000000", "30000", 1).replace("sample1-out", f"f{k}")
            (tmp_path / d / f"f{k}.mci").write_text(t)
    print("case: mcfl CLI, 3 method-1 files in one batch vs alone, -S5", flush=True)
    r = subprocess.run([str(EXE), "f0.mci", "f1.mci", "f2.mci", "-S5", "-F9"], cwd=tmp_path / "many",
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "in one batch" in r.stdout
    for k in range(3):
        r1 = subprocess.run([str(EXE), f"f{k}.mci", "-S5", "-F9"], cwd=tmp_path / "one", capture_output=True,
                            text=True, timeout=300)
        assert r1.returncode == 0, r1.stdout[-3000:] + r1.stderr[-3000:]
        for suffix in ("_Refl.dat", "_Trans.dat", "_time.dat", "_PHD-Ascii.dat"):
            a = (tmp_path / "many" / f"f{k}{suffix}").read_bytes()
            b = (tmp_path / "one" / f"f{k}{suffix}").read_bytes()
            assert a == b, (k, suffix)
