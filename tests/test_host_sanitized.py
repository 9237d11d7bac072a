"""The .mci / bulk-file parser and the output writers (host C++) under AddressSanitizer and
UndefinedBehaviorSanitizer: the reference's sample files and malformed
variants of them (truncated at every line, every line removed or replaced by
a negative, huge, non-numeric or empty value, short / out-of-range / missing
bulk files) must parse or be refused without a memory error, undefined
behaviour or a leak."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "tests", "golden", "ref_inputs")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("asan") / "parse_driver")
    cmd = [cxx, "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-std=c++17",
           "-I" + os.path.join(ROOT, "include"), "-o", out, os.path.join(ROOT, "tests", "sanitize", "parse_driver.cpp"),
           os.path.join(ROOT, "cudamcfl_amd", "csrc", "mcfl_io.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("sanitizer runtime not available: " + r.stderr[-300:])
    return out


def _variants():
    bulk = open(os.path.join(REF, "3dbulk-sample")).read()
    extra = {"short-bulk": bulk[:1000], "bad-bulk": bulk.replace("1", "9", 50)}
    out = []
    for name in ("sample.mci", "sample2.mci"):
        base = open(os.path.join(REF, name)).read().splitlines()
        out.append(base)
        for i in range(len(base)):
            out.append(base[:i])
            out.append(base[:i] + base[i + 1:])
            for rep in ("-1", "99999999999999999999", "abc", "1e40", ""):
                b = list(base)
                b[i] = rep + " # edited"
                out.append(b)
        for bf in ("short-bulk", "bad-bulk", "missing-bulk"):
            out.append([ln if "3dbulk-sample" not in ln else bf for ln in base])
    return out, bulk, extra


def test_parser_under_sanitizers(driver, tmp_path):
    variants, bulk, extra = _variants()
    (tmp_path / "3dbulk-sample").write_text(bulk)
    for k, v in extra.items():
        (tmp_path / k).write_text(v)
    files = []
    for k, v in enumerate(variants):
        p = tmp_path / f"v{k}.mci"
        p.write_text("\n".join(v) + "\n")
        files.append(p.name)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    failures = []
    batches = [files[i:i + 64] for i in range(0, len(files), 64)]  # several files per process
    batches.append(["-w"] + files[::25])  # the writers on a sample of the parsed files
    for b in batches:
        r = subprocess.run([driver] + b, cwd=tmp_path, capture_output=True, text=True, env=env, timeout=300)
        if (r.returncode != 0 or "runtime error" in r.stderr or "AddressSanitizer" in r.stderr
                or "LeakSanitizer" in r.stderr):
            failures.append(r.stderr[-2000:])
    assert not failures, failures[0]
    assert len(files) > 500
