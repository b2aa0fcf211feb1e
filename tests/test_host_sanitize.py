"""The library's host code under AddressSanitizer and UndefinedBehaviorSanitizer
(CPU only, no GPU call): every source of libnekrs_hip is compiled with the
sanitizers on its host side (hipcc -Xarch_host -fsanitize=...; device code as
usual), linked into tests/sanitize/host_driver.cpp, and run on .re2 files that
are valid, truncated, corrupted (counts that do not fit the file, NaN and
huge counts, infinite coordinates, out-of-range curve sides), empty, random
bytes or missing, on valid and malformed multigrid schedules, and through the
FDM setup, GLL, box-mesh, element-length and coarse-id helpers.  Malformed
input must give error codes; any sanitizer report fails the test."""
import concurrent.futures as cf
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
       "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer"]


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    d = tmp_path_factory.mktemp("sanitize")
    csrc = os.path.join(ROOT, "nekrs_amd", "csrc")
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + csrc]
    base = [HIPCC, "--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "-fPIC",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"] + SAN + inc

    def compile_one(src):
        obj = str(d / (os.path.basename(src) + ".o"))
        r = subprocess.run(base + ["-c", src, "-o", obj], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]
        return obj

    srcs = sorted(os.path.join(csrc, f) for f in os.listdir(csrc) if f.endswith(".hip"))
    srcs.append(os.path.join(ROOT, "tests", "sanitize", "host_driver.cpp"))
    with cf.ThreadPoolExecutor(min(8, os.cpu_count() or 4)) as ex:
        objs = list(ex.map(compile_one, srcs))
    exe = str(d / "host_driver")
    r = subprocess.run([HIPCC, "--offload-arch=gfx950"] + SAN + ["-o", exe] + objs,
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe, d


def malformed_files(d):
    import oracle.re2 as r2
    c = r2.box_corners(3, 2, 2)
    good = d / "ok.re2"
    r2.write(good, c, {(0, 0): "W  ", (11, 1): "P  "}, params={(11, 1): (1.0, 4.0)},
             curves=[(0, 1, (0.1, 0.0, 0.0, 0, 0), "m"), (3, 12, (0.5, 0.6, 0.7, 0, 0), "m")])
    data = good.read_bytes()
    E = 12
    off = 84 + 200 * E
    files = {"ok": data, "trunc_header": data[:60], "trunc_elements": data[:300],
             "trunc_tail": data[:-30], "empty": b"",
             "garbage": np.random.default_rng(0).bytes(5000),
             "big_E": b"#v002 999999999 3 999999999 hdr".ljust(80) + data[80:],
             "bad_version": b"#v009" + data[5:]}

    def patched(pos, value):
        b = bytearray(data)
        b[pos:pos + 8] = struct.pack("<d", value)
        return bytes(b)

    files["nan_curves"] = patched(off, float("nan"))
    files["huge_curves"] = patched(off, 1e18)
    files["neg_bc"] = patched(off + 8 + 2 * 64, -5.0)
    files["inf_coordinate"] = patched(84 + 8, float("inf"))
    files["nan_coordinate"] = patched(84 + 16, float("nan"))
    files["curve_side_77"] = patched(off + 16, 77.0)
    files["curve_elem_0"] = patched(off + 8, 0.0)
    files["bc_elem_999"] = patched(off + 8 + 2 * 64 + 8, 999.0)
    files["degenerate"] = patched(84 + 16, struct.unpack("<d", data[84 + 8:84 + 16])[0])
    paths = []
    for name, b in files.items():
        p = d / (name + ".re2")
        p.write_bytes(b)
        paths.append(str(p))
    paths.append(str(d / "missing.re2"))
    return paths


def test_host_code_under_sanitizers(driver):
    exe, d = driver
    paths = malformed_files(d)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe] + paths, capture_output=True, text=True, timeout=600, env=env)
    out = r.stdout + r.stderr
    print(out[-4000:])
    assert "Sanitizer" not in out and "runtime error" not in out, out[-6000:]
    assert r.returncode == 0, out[-3000:]
    assert "host driver done" in out
    # the valid file went all the way through; the broken ones were refused
    ok = out.split("ok.re2:")[1].split(".re2:")[0]
    assert "info rc 0" in ok and "midside N=3 rc 0" in ok
    for bad in ("trunc_elements", "big_E", "nan_curves", "huge_curves", "empty", "missing"):
        assert "%s.re2: info rc -" % bad in out, bad
