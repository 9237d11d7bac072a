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
