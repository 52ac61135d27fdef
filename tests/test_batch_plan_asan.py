"""The batched solver's host-side layout (csrc/batch_plan.h) under AddressSanitizer and UndefinedBehaviorSanitizer
on the CPU, for the sample sets the batch workload and the tests use: tests/native/batch_plan_check.cpp plans
the set for every tile height the autotuner tries and replays every index the batched kernels derive from
the plan (wave tables, AA pull/push slots including link slots, the folded periodic halo, the exchange planes) against the buffers
they address."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

REPO = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
CSRC = os.path.join(REPO, "mf-lbm_amd", "csrc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("asan") / "batch_plan_check")
    src = os.path.join(REPO, "tests", "native", "batch_plan_check.cpp")
    r = subprocess.run([cxx, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                        "-fno-sanitize-recover=all", "-I", CSRC, src, "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def _write_samples(path, samples):
    from mflbm import batch
    with open(path, "wb") as fh:
        fh.write(np.int32(len(samples)).tobytes())
        for s in samples:
            nx, ny, nz = s.shape
            fh.write(np.array([nx, ny, nz, s.jper], np.int32).tobytes())
            fh.write(np.ascontiguousarray(batch.walls_h2(s)).tobytes())


@pytest.mark.parametrize("n,sides", [(16, (48, 64, 96, 128)), (4, (20, 24, 28, 33))])
def test_batch_plan_indices_in_bounds_under_asan(checker, tmp_path, n, sides):
    from mflbm import batch
    samples = batch.rev_samples(n, seed=2024, sides=sides)
    if n == 4:   # the tests' small set also covers y-periodic samples
        samples[1].jper = 1
    path = str(tmp_path / "samples.bin")
    _write_samples(path, samples)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([checker, path], capture_output=True, text=True, env=env, timeout=600)
    sys.stdout.write(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-6000:]
    assert r.stdout.startswith("ok")
