"""The run driver on the GPU path against whole runs of the reference program (the variants of
tests/golden/run_variants.npz: every monitor file line for line, the stop step, the job status and the final
checkpoint bit for bit)."""
import pytest

from mflbm.run import run_directory
from test_run_driver import VARIANTS, RUNS, check_against_reference, prepare

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", VARIANTS)
def test_gpu_run_matches_reference_run(name, tmp_path):
    print(f"[gpu run driver] variant={name}", flush=True)
    run_dir = str(tmp_path / name)
    ctl = prepare(run_dir, name)
    for _ in range(int(RUNS[f"{name}/invocations"])):
        run_directory(run_dir)
    check_against_reference(run_dir, name, ctl, exact=False)
