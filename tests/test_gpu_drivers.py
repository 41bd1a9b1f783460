"""The reference's own driver, gpu_engine_test.cu, compiled unchanged against
this engine's drop-in headers (include/) and libraries by the oracle recipe
(oracle/Makefile -> oracle/_ref/gpu_engine_test), run on the GPU: in one run it
computes y with the reference's own OpenMP CPU sweep (run_cpu_reference_layered,
gpu_engine_test.cu:49-104), checks its 1-thread twin, runs the drop-in
run_gpu_right_multiply_layered / _batched, and exits 0 only when the GPU y
matches the CPU y (|gpu - cpu| / (1 + |cpu|) < 1e-4, gpu_engine_test.cu:251-274).
This tests the drop-in promise that the reference's driver builds and passes
unchanged. The binary is built only where the reference sources are present;
without it the test is skipped."""
import os
import subprocess
import sys

import pytest

import gmr

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DRIVER = os.path.join(ROOT, "oracle", "_ref", "gpu_engine_test")
OWN_DRIVER = os.path.join(ROOT, "g-mm-repair_amd", "bin", "gpu_test")


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    d = tmp_path_factory.mktemp("drv")
    out = []
    for name, m in (("hap", gmr.Matrix.haplotype(20261204, 4096, 2048)),
                    ("rmat", gmr.Matrix.rmat(20261210, 14, 200000, nweights=100))):
        base = str(d / name)
        m.repair().write(base)
        out.append((name, base, m.rows, m.cols))
    return out


@pytest.mark.parametrize("exe", [REF_DRIVER, OWN_DRIVER], ids=["reference_driver", "gpu_test"])
@pytest.mark.parametrize("semiring", ["plustimes", "boolean", "tropical"])
def test_driver_checks_gpu_against_cpu_in_one_run(files, exe, semiring):
    if not os.path.exists(exe):
        pytest.skip(f"{exe} is not built (the reference sources were absent at build time)")
    for name, base, rows, cols in files:
        for B in ("1", "8"):
            env = dict(os.environ, SEMIRING=semiring, GMR_ENERGY_SECONDS="0")
            print(f"{os.path.basename(exe)} {name} {semiring} B={B}", flush=True)
            sys.stdout.flush()
            r = subprocess.run([exe, base, str(rows), str(cols), "1", "repair", B], capture_output=True, text=True,
                               env=env, timeout=120)
            out = r.stdout + r.stderr
            assert r.returncode == 0, out
            assert "CPU OpenMP vs sequential: max_rel_diff=" in r.stdout and "SUCCESS" in r.stdout, out
            assert "SUCCESS: GPU and CPU reference results match" in r.stdout, out
            if B != "1":
                assert "Batched verification: max_rel_diff=" in r.stdout, out
                line = r.stdout.split("Batched verification:")[1].splitlines()[0]
                assert "SUCCESS" in line, line
