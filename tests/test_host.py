"""Host-side C ABI checks that need no GPU: exported symbols, the .mci parser
against hand-derived values of the reference's own sample inputs, writer byte
layouts, the emission plan, and the command-line front end."""
import ctypes as C
import os
import re
import shutil
import subprocess
import sys

import numpy as np
import pytest

import cudamcfl_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_INPUTS = os.path.join(ROOT, "tests", "golden", "ref_inputs")
f32 = np.float32


def header_functions():
    text = open(os.path.join(ROOT, "include", "mcfl.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w \*]*?\b(mcfl_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_function():
    names = header_functions()
    assert len(names) >= 17
    lib = M.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) <= set(M.EXPORTS) | set(names)


@pytest.fixture
def in_ref_inputs(monkeypatch):
    monkeypatch.chdir(REF_INPUTS)


def test_parse_sample2(in_ref_inputs):
    sim = M.Simulation.from_mci("sample2.mci")
    s = sim.s
    assert sim.n_runs == 1 and s.bulk_method == 2
    assert s.outp_filename == b"sample2-out" and s.AorB == b"A"
    assert s.number_of_photons == 500000000 and s.number_of_photons_per_voxel == 40000
    assert s.grid_size == 5 and s.do_fl_sim == 1 and s.fhd_activated == 1
    assert (s.det.dx, s.det.dy) == (f32(0.04), f32(0.04)) and (s.det.nx, s.det.ny) == (250, 250)
    assert (s.xi, s.yi, s.zi) == (-4.0, 0.0, 0.0)
    assert s.do_temp_sim == 1 and (s.det.x_temp_numdets, s.det.y_temp_numdets) == (2, 1)
    assert s.det.temp_det_r == f32(0.5) and s.det.temp_rort == 0
    assert s.det.temp_bins == 1000 and s.det.max_temp == 20000000
    assert s.n_bulks == 2 and s.esp == 5.0
    # grid (4 esp gs)^2 x (esp gs) = 100 x 100 x 25 = 250000 = digits in 3dbulk-sample
    assert (s.nx, s.ny, s.nz) == (100, 100, 25) and s.halfx == 10.0
    b = [s.bulks[i] for i in range(4)]
    assert b[0].n == 1.0 and b[0].mutr == np.finfo(f32).max and b[0].g == 1.0
    assert b[3].n == 1.0 and b[3].mua == 0.0
    # medium 1: mua + flc*AbsCExi = 0.04 + 1e-8*339625 = 0.04339625; muaf = 0.04 + 1e-8*219500
    assert abs(b[1].mua - 0.04339625) < 1e-8 and abs(b[1].muaf - 0.042195) < 1e-8
    assert abs(b[1].mutr - 1 / 48.04) < 1e-8 and b[1].g == f32(0.8) and b[1].eY == f32(0.043)
    assert abs(b[1].albedof - 48 / 48.00339625) < 1e-7
    assert abs(b[2].mua - 0.09698125) < 1e-7 and abs(b[2].muaf - 0.090975) < 1e-7
    assert abs(b[2].mutr - 1 / 48.08) < 1e-8
    # specular start weight for n 1 -> 1.33
    r = ((1.0 - float(f32(1.33))) / (1.0 + float(f32(1.33)))) ** 2
    assert s.start_weight == int(4294967295.0 * (1 - r))
    # labels: the digits of the bulk file, in file order
    digits = np.array([int(c) for c in open("3dbulk-sample").read().split()], np.uint8)
    assert np.array_equal(sim.labels, digits)
    # the _param.dat echo range starts at the output-name line
    raw = open("sample2.mci", "rb").read()
    assert raw[s.begin:].lstrip().startswith(b"sample2-out") and s.end > s.begin


def test_parse_sample_method1(in_ref_inputs):
    sim = M.Simulation.from_mci("sample.mci")
    s = sim.s
    assert s.bulk_method == 1 and s.n_layers == 1
    assert s.number_of_photons_per_voxel == 80000
    L = s.layers[1]
    assert L.z_min == 0.0 and L.z_max == 5.0 and s.esp == 5.0
    assert abs(L.mua - 0.04339625) < 1e-8
    inc = s.inclusion
    assert (inc.x, inc.y, inc.z) == (f32(-1.9), 0.0, f32(1.55)) and inc.r == f32(0.595)
    assert abs(inc.mua - 0.09698125) < 1e-7 and abs(inc.mutr - 1 / 48.08) < 1e-8


def write_mci_variant(tmp_path, **edits):
    src = open(os.path.join(REF_INPUTS, "sample2.mci")).read()
    for old, new in edits.items():
        src = src.replace(old, new, 1)
    p = tmp_path / "v.mci"
    p.write_text(src)
    return str(p)


def test_parser_clamps_photon_counts(tmp_path, monkeypatch):
    monkeypatch.chdir(REF_INPUTS)
    p = write_mci_variant(tmp_path, **{"500000000 ": "1000 ", "40000 ": "32768 "})
    sim = M.Simulation.from_mci(p)
    s = sim.s
    assert s.number_of_photons == 32767 and s.number_of_photons_per_voxel == 32767


@pytest.mark.parametrize("edit", [
    {"1                                                                             # Do fluo": "3 # Do fluo"},
    {"1                                                                             # Do temp": "7 # Do temp"},
    {"0.04\t0.04": "0.04 0.04 0.04"},
])
def test_parser_rejects_malformed(tmp_path, monkeypatch, edit):
    monkeypatch.chdir(REF_INPUTS)
    p = write_mci_variant(tmp_path, **edit)
    with pytest.raises(M.McflError):
        M.Simulation.from_mci(p)


def test_parser_missing_file():
    with pytest.raises(M.McflError):
        M.Simulation.from_mci("/nonexistent/file.mci")


def test_dims_extension_roundtrip(tmp_path, monkeypatch):
    from cudamcfl_amd.media import write_bulk, write_mci, inclusion_medium
    monkeypatch.chdir(tmp_path)
    lab, mats = inclusion_medium(seed=5, shape=(8, 12, 10))
    write_bulk("b.txt", lab)
    write_mci("m.mci", "out", "b.txt", mats, esp=0.4, grid_size=20, photons=10 ** 6, dims=(10, 12, 8))
    s = M.Simulation.from_mci("m.mci")
    assert (s.s.nx, s.s.ny, s.s.nz) == (10, 12, 8)
    assert s.s.halfx == f32(10) / f32(40) and s.s.halfy == f32(12) / f32(40)
    assert np.array_equal(s.labels, lab.reshape(-1))
    b1 = s.s.bulks[1]
    assert b1.n == f32(mats[0][0]) and b1.g == f32(mats[0][7])


def test_interpret_arg():
    lib = M.lib()
    argv = (C.c_char_p * 4)(b"mcfl", b"x.mci", b"-S12345", b"-A")
    seed, ign = C.c_uint64(0), C.c_int(0)
    assert lib.mcfl_interpret_arg(4, argv, C.byref(seed), C.byref(ign)) == 0
    assert seed.value == 12345 and ign.value == 1
    argv = (C.c_char_p * 3)(b"mcfl", b"x.mci", b"-Q")
    assert lib.mcfl_interpret_arg(3, argv, C.byref(seed), C.byref(ign)) == 1


def test_writers_byte_layout(tmp_path, monkeypatch):
    monkeypatch.chdir(REF_INPUTS)
    sim = M.Simulation.from_mci("sample2.mci")
    s = sim.s
    s.outp_filename = str(tmp_path / "w").encode()
    s.det.nx, s.det.ny = 3, 2
    s.do_temp_sim = 1
    s.det.temp_bins = 2
    npix = 3 * 3
    r = M.ExcitationResult(sim, fhd=False)
    r.Rd[:] = np.arange(npix, dtype=np.uint64) * 1000003
    r.Tt[:] = np.arange(npix, dtype=np.uint64)[::-1] * 77
    r.time[:4] = [5, 6, 7, 8]
    assert M.lib().mcfl_write_simulation_results(C.byref(s), C.byref(r.t), 1.25) == 0
    scale = float(0xFFFFFFFF) * 500000000.0
    scale2 = scale * float(f32(0.04)) * float(f32(0.04))

    def img(a):
        return "".join("".join(" %E " % (float(a[y * 3 + x]) / scale2) for x in range(3)) + " \n " for y in range(2))

    assert open(tmp_path / "w_Refl.dat").read() == img(r.Rd)
    assert open(tmp_path / "w_Trans.dat").read() == img(r.Tt)
    t = open(tmp_path / "w_time.dat").read()
    assert t.startswith("#x y timebins[numbins]\n")
    assert "%f %f " % (-0.5, 0.0) + " %E " % (5 / scale) + " %E " % (7 / scale) + " \n " in t
    p = open(tmp_path / "w_param.dat", "rb").read()
    # the header of CUDAMCFLio.cu:100-112 byte for byte (after the build-date line)
    head = (b"A1 \t# Version number of the file format.\n\n####\n# Data categories include: \n"
            b"# InParm, RAT, \n# A_l, A_z, Rd_r, Rd_a, Tt_r, Tt_a, \n# A_rz, Rd_ra, Tt_ra \n####\n\n"
            b"# User time: 1.25 sec\n\nInParam\t\t# Input parameters:\n")
    first, rest = p.split(b"\n\n", 1)
    assert first.startswith(b"Compilation date: ") and first.endswith(b". ")
    assert rest.startswith(head)
    raw = open("sample2.mci", "rb").read()
    assert p.split(b"InParam\t\t# Input parameters:\n", 1)[1] == raw[s.begin:s.end]


def test_write_phd_and_fl(tmp_path):
    lab = np.ones((2, 2, 3), np.uint8)
    sim = M.Simulation.voxel(lab, [[1.4, 0.1, 10, 0, 0, 0, 0, 0.9]], esp=0.2, grid_size=10,
                             det=(0.1, 0.1, 2, 2, 0, 0), outp=str(tmp_path / "p"))
    Fx = np.arange(12, dtype=np.float64) * 1.5e-3
    assert M.lib().mcfl_write_phd(C.byref(sim.s), Fx.ctypes.data_as(M.F64P)) == 0
    txt = open(tmp_path / "p_PHD-Ascii.dat").read()
    assert txt == "3\t2\t2\n" + "".join("%.10E\n" % v for v in Fx)
    R = np.array([1.0, 2.5e-7, 3.0, 0.0])
    T = np.array([0.5, 0.25, 0.125, 1e-12])
    assert M.lib().mcfl_write_fl(C.byref(sim.s), R.ctypes.data_as(M.F64P), T.ctypes.data_as(M.F64P)) == 0
    assert open(tmp_path / "p_FlR.dat").read() == " %E  %E  \n  %E  %E  \n " % tuple(R)
    assert open(tmp_path / "p_FlT.dat").read() == " %E  %E  \n  %E  %E  \n " % tuple(T)


def test_fl_plan_matches_reference_weights():
    import media
    sim = media.small_sim(photons=1000)
    rng = np.random.default_rng(0)
    Fx = rng.uniform(0, 1e-3, sim.nvox)
    n = 50000
    vox, scale, unit = M.fl_plan(sim, Fx, n)
    lab = sim.labels
    bul = [sim.s.bulks[i] for i in range(sim.s.n_bulks + 2)]
    exp_vox = [v for v in range(sim.nvox) if bul[lab[v]].albedof < 1]
    assert np.array_equal(vox, exp_vox)
    voxelw = np.array([bul[lab[v]].eY * float(f32(1) - f32(bul[lab[v]].albedof)) * Fx[v] / (n * 4294967295.0)
                       for v in vox])
    np.testing.assert_allclose(scale * unit, voxelw, rtol=1e-14)
    assert unit > 0 and (voxelw * n * 2.0 ** 32).sum() / unit <= 2.0 ** 62 * (1 + 1e-12)


def test_cli_without_arguments():
    exe = os.path.join(M.BUILD_DIR, "mcfl")
    p = subprocess.run([exe], capture_output=True, text=True)
    assert p.returncode == 1 and "Not enough input arguments!" in p.stdout


def test_cli_flag_in_place_of_the_input_file():
    """A flag where the .mci file belongs is read as the file name, as the
    reference does (CUDAMCFLmain.cu:305-320): the parser fails and the exit
    status is 1, before any GPU work."""
    exe = os.path.join(M.BUILD_DIR, "mcfl")
    p = subprocess.run([exe, "-S5"], capture_output=True, text=True)
    assert p.returncode == 1 and "Something wrong with read_simulation_data!" in p.stdout
    assert "No usable GPU" not in p.stdout


def test_cli_rejects_unknown_argument(monkeypatch):
    monkeypatch.chdir(REF_INPUTS)
    exe = os.path.join(M.BUILD_DIR, "mcfl")
    p = subprocess.run([exe, "sample2.mci", "-Z"], capture_output=True, text=True)
    assert p.returncode == 1 and "Unknown argument -Z!" in p.stdout


def test_output_file_readers_round_trip(tmp_path):
    """Files written by the C ABI's writers read back to the written values."""
    from cudamcfl_amd.files import read_image, read_phd, read_time
    sim = M.Simulation.voxel(np.ones((3, 4, 5), np.uint8), [[1.4, 0.1, 10.0, 0, 0, 0, 0, 0.9]], esp=0.3,
                             grid_size=10, det=(0.1, 0.2, 5, 4, 0.0, 0.0), photons=1000,
                             temporal=(0.0, 0.0, 2, 1, 0.5, 0.5, 0.3, 0, 3, 1000), outp=str(tmp_path / "rt"))
    s = sim.s
    src = tmp_path / "rt.mci"
    src.write_text("# input echo\n")
    s.inp_filename = str(src).encode()
    s.begin, s.end = 0, 0
    rng = np.random.default_rng(5)
    Fx = rng.random(sim.nvox)
    assert M.lib().mcfl_write_phd(C.byref(s), Fx.ctypes.data_as(M.F64P)) == 0
    got = read_phd(tmp_path / "rt_PHD-Ascii.dat")
    assert got.shape == (3, 4, 5)
    assert np.allclose(got.reshape(-1), Fx, rtol=1e-10, atol=0)
    assert read_phd(tmp_path / "rt_PHD-Ascii.dat", reference_axes=True).shape == (5, 4, 3)
    r = M.ExcitationResult(sim, fhd=False)
    r.Rd[:] = rng.integers(0, 2 ** 40, r.Rd.size, dtype=np.uint64)
    r.Tt[:] = rng.integers(0, 2 ** 40, r.Tt.size, dtype=np.uint64)
    r.time[:] = rng.integers(0, 2 ** 40, r.time.size, dtype=np.uint64)
    assert M.lib().mcfl_write_simulation_results(C.byref(s), C.byref(r.t), 0.5) == 0
    scale = float(0xFFFFFFFF) * 1000.0
    area = float(np.float32(0.1)) * float(np.float32(0.2))
    R = read_image(tmp_path / "rt_Refl.dat")
    assert R.shape == (4, 5)
    assert np.allclose(R.reshape(-1), r.Rd[:20].astype(np.float64) / (scale * area), rtol=1e-6)
    pos, hist = read_time(tmp_path / "rt_time.dat")
    assert pos.shape == (2, 2) and hist.shape == (2, 3)
    assert np.allclose(hist[1], r.time[[1, 3, 5]].astype(np.float64) / scale, rtol=1e-6)


def test_reference_shaped_sphere_problem():
    """80x80x20 voxels of 1 mm from esp = 2 cm and grid_size 10 (CUDAMCFLio.cu:609-612),
    a voxelized sphere of radius 5 voxels, the parser's derived optics."""
    import numpy as np
    from cudamcfl_amd.media import sphere_simulation
    sim = sphere_simulation(photons=1000)
    s = sim.s
    assert (s.nx, s.ny, s.nz) == (80, 80, 20) and abs(s.halfx - 4.0) < 1e-6
    counts = np.bincount(sim.labels.ravel(), minlength=3)
    assert abs(counts[2] - 4.0 / 3.0 * np.pi * 5 ** 3) < 0.1 * 4.0 / 3.0 * np.pi * 5 ** 3
    b = s.bulks
    assert abs(b[1].mua - 0.11) < 1e-6 and abs(b[2].mua - 0.6) < 1e-6        # mua + flc AbsCExi
    assert abs(b[1].muaf - 0.08) < 1e-6 and abs(b[2].muaf - 0.08) < 1e-6      # mua + flc AbsCEmi
    assert abs(b[1].mutr - 1.0 / 100.1) < 1e-7 and abs(b[2].mutr - 1.0 / 100.5) < 1e-7


def test_entry_points_refuse_null_arguments():
    """Every entry point that takes a context, batch or server handle returns
    MCFL_ERR_ARG (or 0 photons) for a null handle or a null required pointer,
    before touching the device."""
    lib = M.lib()
    M._bind_engine(lib)
    ERR_ARG = 1
    sim = M.Simulation.voxel(np.ones((2, 2, 2), np.uint8), [(1.4, 1.0, 100.0, 0.0, 0.0, 0.0, 0.1, 0.9)],
                             esp=0.2, grid_size=10, det=(0.1, 0.1, 4, 4, 0.0, 0.0), photons=10)
    S = C.byref(sim.s)
    Fx = np.zeros(8)
    F = Fx.ctypes.data_as(M.F64P)
    rng = M.Rng()
    tune = M.Tune()
    calls = {
        "mcfl_ctx_create": lambda: lib.mcfl_ctx_create(0, None, None),
        "mcfl_ctx_info": lambda: lib.mcfl_ctx_info(None, None, None, None, 0),
        "mcfl_run_excitation": lambda: lib.mcfl_run_excitation(None, S, C.byref(rng), None),
        "mcfl_run_fluorescence": lambda: lib.mcfl_run_fluorescence(None, S, F, 10, 1, F, F, None),
        "mcfl_run_fluorescence_lanes": lambda: lib.mcfl_run_fluorescence_lanes(None, S, F, 10, None, F, F, None),
        "mcfl_fl_plan": lambda: lib.mcfl_fl_plan(None, F, 10, None, None, None, None),
        "mcfl_batch_create": lambda: lib.mcfl_batch_create(None, None, 0, None, 0, None),
        "mcfl_batch_run": lambda: lib.mcfl_batch_run(None, None, None),
        "mcfl_batch_autotune": lambda: lib.mcfl_batch_autotune(None, 0.1, None),
        "mcfl_batch_configure": lambda: lib.mcfl_batch_configure(None, C.byref(tune)),
        "mcfl_run_batch": lambda: lib.mcfl_run_batch(None, None, 0, None, 0, None),
        "mcfl_server_create": lambda: lib.mcfl_server_create(None, 0, 1, None),
        "mcfl_server_add_medium": lambda: lib.mcfl_server_add_medium(None, S, None),
        "mcfl_server_submit": lambda: lib.mcfl_server_submit(None, None, None),
        "mcfl_server_step": lambda: lib.mcfl_server_step(None, None, 0, None, None),
        "mcfl_server_collect": lambda: lib.mcfl_server_collect(None, 0, None),
        "mcfl_server_stats": lambda: lib.mcfl_server_stats(None, None, None, None),
        "mcfl_server_counters": lambda: lib.mcfl_server_counters(None, None, None, None, None),
    }
    for name, call in calls.items():
        assert call() == ERR_ARG, name
    assert lib.mcfl_DoOneSimulation(None, S, None, None, F) == 0
    assert lib.mcfl_DoOneSimulationFl(None, S, None, None, None, None) == 0
    assert lib.mcfl_run_excitation_noalloc(None, S, None, None) == ERR_ARG
    lib.mcfl_ctx_destroy(None)
    lib.mcfl_batch_destroy(None)
    lib.mcfl_server_destroy(None)


def test_integration_example_compiles_and_links(tmp_path):
    """The C snippet that shows a maintainer how to replace the reference's
    main() compiles as C against include/mcfl.h and links against libmcfl.so."""
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    code = re.search(r"```c\n(.*?)```", text, re.S).group(1)
    body = "\n".join("  " + ln for ln in code.splitlines() if not ln.startswith("#include"))
    src = ("#include <stdlib.h>\n#include <time.h>\n#include \"mcfl.h\"\n"
           "int main(int argc, char** argv) {\n  uint64_t seed = 1;\n" + body + "\n  (void)done;\n  return 0;\n}\n")
    (tmp_path / "main.c").write_text(src)
    build = os.path.dirname(M.LIB_PATH)
    r = subprocess.run(["gcc", "-O2", "-Wall", "-Werror", "-o", str(tmp_path / "cuda_fl"), str(tmp_path / "main.c"),
                        "-I" + os.path.join(ROOT, "include"), "-L" + build, "-lmcfl", "-Wl,-rpath," + build],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_no_kernel_variant_spills(tmp_path):
    """Every transport kernel variant fits its vector register budget without
    spilling to scratch memory (compiled for gfx950 here; the counts come
    from the code object metadata of the device assembly).  Scalar register
    spills, which go to lanes of a vector register, are allowed."""
    out = str(tmp_path)
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "cudamcfl_amd", "csrc"), f"OUT={out}", "asm"], check=True,
                   capture_output=True, timeout=600)
    text = open(os.path.join(out, "mcfl_transport.s")).read()
    kernels = re.findall(r"\.name:\s+(_Z\w*mcfl_transport_kernel\w*)(.*?)\.vgpr_spill_count:\s+(\d+)", text, flags=re.S)
    assert len(kernels) >= 20
    spills = {name: int(n) for name, _, n in kernels if int(n) != 0}
    assert not spills, spills


def test_volume_tally_kernel_choice(tmp_path):
    """A job launch runs the kernel compiled with the volume tallies its jobs
    carry: A_xyz (with or without fhd) needs the full kernel, fhd alone the
    fhd kernel, neither (whatever the other flags) the kernel without them
    (mcfl_vol_of in the kernels' shared header, compiled here with g++)."""
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src = tmp_path / "vol.cpp"
    src.write_text('''#include <stdio.h>
#include "mcfl_device.h"
int main() {
  const unsigned f[] = {0u, MCFL_F_FHD, MCFL_F_AXYZ, MCFL_F_FHD | MCFL_F_AXYZ, MCFL_F_SQ | MCFL_F_TEMP | MCFL_F_SCALED,
                        MCFL_F_FHD | MCFL_F_SQ, MCFL_F_AXYZ | MCFL_F_TEMP};
  for (unsigned k = 0; k < sizeof f / sizeof f[0]; ++k) printf("%u ", mcfl_vol_of(f[k]));
  printf("%d %d %d\\n", (int)MCFL_VOL_ALL, (int)MCFL_VOL_FHD, (int)MCFL_VOL_NONE);
  return 0;
}
''')
    exe = str(tmp_path / "vol")
    r = subprocess.run(["g++", "-std=c++17", "-o", exe, str(src), "-I" + os.path.join(ROOT, "cudamcfl_amd", "csrc"),
                        "-I" + os.path.join(ROOT, "include")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    v = [int(x) for x in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    ALL, FHD, NONE = v[-3:]
    assert v[:-3] == [NONE, FHD, ALL, ALL, NONE, FHD, ALL]
    assert ALL == 0  # the zero-initialised launch arguments compile every tally in


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; --gpus N above 1 exits with a message before
    touching a device."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2 and "one GPU" in (r.stdout + r.stderr)


def test_python_mirror_matches_the_c_layout(tmp_path):
    """Every ctypes structure of the Python binding has the size and field
    offsets of its C counterpart in include/mcfl.h (compiled here with gcc)."""
    import shutil
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    pairs = [("mcfl_layer", M.Layer), ("mcfl_bulk", M.Bulk), ("mcfl_inclusion", M.Inclusion), ("mcfl_det", M.Det),
             ("mcfl_sim", M.Sim), ("mcfl_rng", M.Rng), ("mcfl_tallies", M.Tallies), ("mcfl_fl_result", M.FlResult),
             ("mcfl_job", M.Job), ("mcfl_request", M.Request), ("mcfl_request_result", M.RequestResult),
             ("mcfl_tune", M.Tune)]
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "mcfl.h"', "int main(void) {"]
    expect = []
    for cname, py in pairs:
        lines.append(f'  printf("%zu\\n", sizeof({cname}));')
        expect.append(C.sizeof(py))
        for f in py._fields_:
            lines.append(f'  printf("%zu\\n", offsetof({cname}, {f[0]}));')
            expect.append(getattr(py, f[0]).offset)
    lines += ["  return 0;", "}"]
    (tmp_path / "layout.c").write_text("\n".join(lines) + "\n")
    exe = str(tmp_path / "layout")
    r = subprocess.run(["gcc", "-o", exe, str(tmp_path / "layout.c"), "-I" + os.path.join(ROOT, "include")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    got = [int(v) for v in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    assert got == expect
