"""The host code built with AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

tests/sanitize_driver.cpp drives the seeded generator, the PNG codec and frame loader, the CPU
oracle (a whole 60x80 solve and every stage entry point) and the drop-in CSF_cuda class on its
failure paths; the pdflow-impair command line is built the same way and run on its argument and
file errors.  Every sanitizer report is fatal (-fno-sanitize-recover, halt_on_error), so the test
fails on the first out-of-bounds access, use after free or undefined operation.  The HIP kernels
are not part of this build: GPU code has no sanitizer on this hardware.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "pd-flow_amd", "csrc")
PKG = os.path.join(ROOT, "pd-flow_amd", "pdflow")
INC = os.path.join(ROOT, "include")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=0",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="2")


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    return r.returncode, r.stdout + r.stderr


@pytest.fixture(scope="module")
def build_dir(tmp_path_factory):
    if not shutil.which("g++") or not os.path.exists(os.path.join(PKG, "libpdflow_hip.so")):
        pytest.fail("g++ and the built libraries are needed (__graft_entry__.build())")
    d = tmp_path_factory.mktemp("asan")
    objs = []
    for src, std in ((os.path.join(ROOT, "oracle", "pdflow_oracle.c"), "-std=c11"),
                     (os.path.join(CSRC, "synth.c"), "-std=gnu11"),
                     (os.path.join(CSRC, "pdflow_png.c"), "-std=gnu11")):
        o = str(d / (os.path.basename(src) + ".o"))
        rc, out = _run(["gcc", *SAN, std, "-ffp-contract=off", "-fopenmp", f"-I{INC}", "-c", src, "-o", o])
        assert rc == 0, out
        objs.append(o)
    for src in (os.path.join(CSRC, "csf_cuda.cpp"), os.path.join(ROOT, "tests", "sanitize_driver.cpp")):
        o = str(d / (os.path.basename(src) + ".o"))
        rc, out = _run(["g++", *SAN, "-std=c++17", f"-I{INC}", "-c", src, "-o", o])
        assert rc == 0, out
        objs.append(o)
    exe = str(d / "driver")
    rc, out = _run(["g++", *SAN, "-fopenmp", "-o", exe, *objs, f"-L{PKG}", "-lpdflow_hip", "-lz", "-lm",
                    f"-Wl,-rpath,{PKG}"])
    assert rc == 0, out
    impair = str(d / "pdflow-impair")
    png_o = objs[2]
    rc, out = _run(["g++", *SAN, "-std=c++17", f"-I{INC}", "-o", impair, os.path.join(CSRC, "pdflow_impair.cpp"),
                    png_o, objs[3], f"-L{PKG}", "-lpdflow_hip", "-lz", f"-Wl,-rpath,{PKG}"])
    assert rc == 0, out
    return d


def test_host_code_is_clean_under_asan_and_ubsan(build_dir):
    rc, out = _run([str(build_dir / "driver"), str(build_dir)], env=ENV, timeout=600)
    assert rc == 0 and "sanitized host code ok" in out, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]


def test_command_line_errors_are_clean_under_asan_and_ubsan(build_dir):
    exe = str(build_dir / "pdflow-impair")
    missing = str(build_dir / "none.png")
    for args in (["--help"], ["--rows", "100"], ["--rows"], ["--bogus"],
                 ["--rows", "120", "--i1", missing, "--i2", missing, "--z1", missing, "--z2", missing,
                  "--out", str(build_dir / "r"), "--no-show"]):
        rc, out = _run([exe, *args], env=ENV, timeout=120)
        assert "AddressSanitizer" not in out and "runtime error" not in out, (args, out[-3000:])
        assert rc in (0, 1), (args, rc, out[-3000:])
