"""The host-side C and C++ code built with AddressSanitizer and
UndefinedBehaviorSanitizer (g++/gcc, no GPU), every report fatal:

* the library's host code (csrc/ps_host.cpp, csrc/ps_learned.cpp) through
  tests/sanitize/host_harness.cpp: the reference's sample input, the ten
  committed work-unit inputs and a 5,000-point synthetic input, each read
  with every reader option and as truncated copies; the row text, the rows
  reader, the three checkpoint formats and damaged checkpoints, the pole
  reduction with NaN fits, and the network-file reader on the shipped
  network and on damaged copies;
* the oracle's driver (oracle/ps_oracle_main.c + ps_oracle.c): the
  convexity-weight precalc and the scan of one trial period of a small
  synthetic asteroid.
"""
import os
import shutil
import subprocess

import pytest

from periodsearch_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SAN = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
       "-fno-builtin-sin", "-fno-builtin-cos", "-fno-builtin-sincos"]
ENV = dict(os.environ, ASAN_OPTIONS="halt_on_error=1:detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1")


def _sanitizers_available(tmp_path):
    src = tmp_path / "t.c"
    src.write_text("int main(void) { return 0; }\n")
    r = subprocess.run(["gcc", *SAN, "-o", str(tmp_path / "t"), str(src)], capture_output=True, text=True)
    return r.returncode == 0 and subprocess.run([str(tmp_path / "t")], env=ENV).returncode == 0


@pytest.fixture
def san(tmp_path):
    if not shutil.which("gcc") or not shutil.which("g++") or not _sanitizers_available(tmp_path):
        pytest.skip("gcc/g++ with ASan and UBSan runtimes are not available")
    return tmp_path


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, env=ENV, **kw)
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0, r.stderr[-4000:]
    return r


def test_host_library_code_under_sanitizers(san):
    exe = str(san / "host_harness")
    _run(["g++", *SAN, "-std=c++17", "-pthread", "-I", os.path.join(ROOT, "include"),
          "-I", os.path.join(ROOT, "periodsearch_amd", "csrc"), "-o", exe,
          os.path.join(HERE, "sanitize", "host_harness.cpp"),
          os.path.join(ROOT, "periodsearch_amd", "csrc", "ps_host.cpp"),
          os.path.join(ROOT, "periodsearch_amd", "csrc", "ps_learned.cpp")], timeout=300)
    big = str(san / "big_in")
    synth.write_workload("big", big, per_start=7.0, per_end=7.6)
    wu = sorted(os.path.join(HERE, "golden", "wu", f) for f in os.listdir(os.path.join(HERE, "golden", "wu")))
    scratch = san / "scratch"
    scratch.mkdir()
    _run([exe, str(scratch), os.path.join(ROOT, "periodsearch_amd", "proposal_net.npz"),
          os.path.join(HERE, "golden", "ps_sample_out"), os.path.join(HERE, "golden", "ps_sample_in"), *wu, big],
         timeout=300)


def test_oracle_under_sanitizers(san):
    exe = str(san / "orc")
    _run(["gcc", *SAN, "-I", os.path.join(ROOT, "oracle"), "-o", exe,
          os.path.join(ROOT, "oracle", "ps_oracle_main.c"), os.path.join(ROOT, "oracle", "ps_oracle.c"), "-lm"],
         timeout=300)
    cfg = {k: v for k, v in synth.WORKLOADS["small"].items() if k not in ("seed", "per_start", "per_end")}
    cfg.update(lmax=3, mmax=3, incl_deg=10.0)
    ast = synth.make_asteroid(0, **cfg)
    inp = str(san / "small_in")
    # one trial period; a starting weight that ends the precalc after one round
    synth.write_input(inp, ast, 7.0, 7.0, conw=9.0)
    out = str(san / "small_out")
    _run([exe, inp, out], timeout=600)
    rows = open(out).read().split("\n")
    assert len([r for r in rows if r.strip()]) == 1
