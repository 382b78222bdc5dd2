"""Host-side sanitizer build of the C ABI (CPU only).

The host code of csg_capi.hip, csg_batch.hip and csg_noalloc.hip (argument
validation, context and ensemble set-up and clean-up, twiddle generation) is
compiled with AddressSanitizer and UndefinedBehaviorSanitizer (host pass
only: `-Xarch_host` before each `-fsanitize=`; the GPU code objects are
built as usual), julia_rng.cpp likewise, and linked with the kernel-table
units of the regular build into tests/cpp/host_sanitize_check.cpp.  The
driver runs the RNG and initial_fields over edge sizes (compared here with
the oracle) and every entry point's argument rejection and no-device
clean-up path; any sanitizer report fails the test.
"""
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import core_shell_oracle as o
from oracle import julia_rng

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "embedded-spinodal-core-shell-julia_amd")
CSRC = os.path.join(PKG, "csrc")
HIPCC = "/opt/rocm/bin/hipcc"
CLANGXX = "/opt/rocm/llvm/bin/clang++"
SAN = ["address", "undefined"]
HOST_SAN = [f for s in SAN for f in ("-Xarch_host", f"-fsanitize={s}")] + ["-Xarch_host", "-fno-omit-frame-pointer"]
SANITIZED_UNITS = ("csg_capi.hip", "csg_batch.hip", "csg_noalloc.hip")


def _run(cmd, **kw):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, **kw)
    assert r.returncode == 0, " ".join(cmd) + "\n" + r.stdout[-4000:]
    return r.stdout


def _kernel_table_objects():
    import importlib.util
    spec = importlib.util.spec_from_file_location("csg_build", os.path.join(PKG, "build.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    objs = [os.path.join(b.BUILD, os.path.splitext(src)[0] + ".o") for src, _ in b.UNITS
            if src not in SANITIZED_UNITS]
    if not all(os.path.exists(p) for p in objs):
        b.build()
    return objs, b.HIP_FLAGS


def test_host_code_under_asan_ubsan(tmp_path):
    objs, hip_flags = _kernel_table_objects()
    flags = [f for f in hip_flags if f != "-O3"] + ["-O1", "-g"]
    jobs = []
    san_objs = []
    for src in SANITIZED_UNITS:
        obj = str(tmp_path / (os.path.splitext(src)[0] + ".o"))
        jobs.append([HIPCC, *flags, *HOST_SAN, "-c", os.path.join(CSRC, src), "-o", obj])
        san_objs.append(obj)
    rng_obj = str(tmp_path / "julia_rng.o")
    jobs.append([CLANGXX, "-O1", "-g", "-std=c++17", "-fPIC", "-fsanitize=address,undefined",
                 "-fno-omit-frame-pointer", "-c", os.path.join(CSRC, "julia_rng.cpp"), "-o", rng_obj])
    with ThreadPoolExecutor(max_workers=4) as ex:
        list(ex.map(lambda c: _run(c, timeout=900), jobs))
    exe = str(tmp_path / "host_sanitize_check")
    drv = str(tmp_path / "driver.o")
    _run([HIPCC, "-O1", "-g", "-std=c++17", *HOST_SAN, "-c", os.path.join(ROOT, "tests", "cpp", "host_sanitize_check.cpp"),
          "-o", drv], timeout=300)
    _run([CLANGXX, "-fsanitize=address,undefined", "-o", exe, drv, *san_objs, rng_obj, *objs,
          "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-lamdhip64"], timeout=300)
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=0:verify_asan_link_order=0:abort_on_error=0:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    env["HIP_VISIBLE_DEVICES"] = ""   # this check never touches a device
    data = str(tmp_path / "rng.bin")
    r = subprocess.run([exe, data], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600, env=env)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout[-6000:]
    assert "runtime error" not in r.stdout and "AddressSanitizer" not in r.stdout, r.stdout[-6000:]
    # the driver's RNG and initial_fields outputs against the oracle
    got = np.fromfile(data, dtype=np.float64)
    k = 0
    for seed in (5749, 0, 1, 123456789012):
        for count in (1, 2, 7, 382, 383, 384, 385, 1000, 1001, 1002, 1003, 4096):
            ref = julia_rng.MersenneTwister(seed).rand(count)
            assert np.array_equal(got[k:k + count], ref), (seed, count)
            k += count
    for n in (1, 2, 3, 16, 17, 64):
        c, phi = o.initial_fields(n, radius=0.3 * n, width=2.0, seed=5749)
        c_got = got[k:k + n * n].reshape((n, n), order="F")
        k += n * n
        phi_got = got[k:k + n * n].reshape((n, n), order="F")
        k += n * n
        # libm tanh vs numpy tanh may differ in the last ulp (as in test_host)
        assert np.max(np.abs(phi_got - phi)) <= 2.3e-16, n
        assert np.array_equal(phi_got >= 0.5, phi >= 0.5), n
        assert np.max(np.abs(c_got - c)) <= 1e-17, n
    assert k == got.size
