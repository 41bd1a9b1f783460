"""Command-line tools: host_scheduler against the reference's own inspector,
and gpu_test's error paths (no GPU needed for these)."""
import os
import subprocess

import numpy as np
import pytest

import gmr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "g-mm-repair_amd", "bin")
REF_HS = os.path.join(ROOT, "oracle", "_ref", "host_scheduler")


def _grammar(tmp_path):
    m = gmr.Matrix.haplotype(41, 600, 256, founders=10, seg_mean=24.0)
    g = m.repair()
    base = str(tmp_path / "h")
    g.write(base)
    return m, g, base


def test_host_scheduler_matches_the_reference_inspector(tmp_path):
    m, g, base = _grammar(tmp_path)
    mine = subprocess.run([os.path.join(BIN, "host_scheduler"), base], capture_output=True, text=True)
    assert mine.returncode == 0
    if os.path.exists(REF_HS):
        ref = subprocess.run([REF_HS, base], capture_output=True, text=True)
        assert ref.returncode == 0
        assert mine.stdout == ref.stdout
    full = subprocess.run([os.path.join(BIN, "host_scheduler"), base, "600", "256"], capture_output=True, text=True)
    assert full.returncode == 0 and "[Wave64]" in full.stdout and "[Resident]" in full.stdout
    info = g.schedule(600, 256, "refcompat").info()
    assert f"+pt: {info['num_passthrough']}" in full.stdout


def test_gpu_test_argument_and_input_errors(tmp_path):
    exe = os.path.join(BIN, "gpu_test")
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage" in r.stderr
    r = subprocess.run([exe, str(tmp_path / "missing"), "10", "10"], capture_output=True, text=True)
    assert r.returncode == 1
    m, g, base = _grammar(tmp_path)
    bad = tmp_path / "x.bin"
    np.zeros(255, np.float32).tofile(bad)   # cols is 256
    r = subprocess.run([exe, base, "600", "256"], capture_output=True, text=True, env=dict(os.environ, XVEC=str(bad)))
    assert r.returncode == 1 and "expected cols=256" in r.stderr


def test_rocsparse_test_argument_errors(tmp_path):
    """The vendor driver rejects what the reference's cusparse_test rejects,
    before any GPU call."""
    exe = os.path.join(BIN, "rocsparse_test")
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage" in r.stderr
    r = subprocess.run([exe, str(tmp_path / "missing"), "10", "10"], capture_output=True, text=True)
    assert r.returncode == 1 and "Error" in r.stderr
    m, g, base = _grammar(tmp_path)
    r = subprocess.run([exe, base, "600", "256", "5", "recompress"], capture_output=True, text=True)
    assert r.returncode == 1 and "no longer supported" in r.stderr
    r = subprocess.run([exe, base, "599", "256"], capture_output=True, text=True)  # one row too few
    assert r.returncode == 1 and "more than 599 rows" in r.stderr


def test_bench_measures_one_gpu_only():
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_level_times_refuses_without_a_gpu():
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from gmr import level_times
    with pytest.raises(SystemExit, match="needs a GPU"):
        level_times.main(["--rows", "64", "--cols", "32"])
