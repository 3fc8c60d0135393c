"""The FP64 oracle under AddressSanitizer and UndefinedBehaviorSanitizer (no GPU).

oracle/sanitize_main.c drives oracle/xfdtd_oracle.c through five small scenes
that together make every builder and read-back call: CPML with kappa > 1,
oblique TFSF3D, NffftFd and its far field, Drude/Debye/Lorentz media, every
lumped element kind, ObjectPlane, line monitors, the adjoint gradient, the
broadband surface and a 2-D TMz grid with TFSF2D.  The Makefile's `sanitize`
target builds it with -fsanitize=address,undefined -fno-sanitize-recover=all,
so any out-of-bounds access, use after free, leak (LeakSanitizer) or
undefined arithmetic in those loops ends the run with a report and a non-zero
status.  OpenMP runs with two threads so the parallel loops are covered too.
"""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


def test_oracle_scenes_clean_under_asan_and_ubsan():
    if shutil.which("gcc") is None or shutil.which("make") is None:
        pytest.skip("gcc and make are needed to build the sanitized oracle")
    build = subprocess.run(["make", "-C", ORACLE, "sanitize"], capture_output=True, text=True, timeout=300)
    if build.returncode != 0 and ("lasan" in build.stderr or "lubsan" in build.stderr):
        pytest.skip("the sanitizer runtimes are not installed: " + build.stderr[-300:])
    assert build.returncode == 0, build.stderr[-2000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="2")
    run = subprocess.run([os.path.join(ORACLE, "_san", "oracle_san")], capture_output=True, text=True,
                         timeout=600, env=env, cwd=ORACLE)
    report = run.stdout[-2000:] + run.stderr[-4000:]
    assert run.returncode == 0, report
    assert "AddressSanitizer" not in run.stderr and "runtime error" not in run.stderr, report
    assert run.stdout.strip().endswith("sanitized oracle scenes: ok"), report
    for scene in ("scatter", "media", "adjoint", "broadband", "tmz"):
        assert scene + ":" in run.stdout, report
