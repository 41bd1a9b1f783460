"""The vendor baselines on the GPU: rocSPARSE SpMV / SpMM of the uncompressed
matrix against the oracle's brute force (every CSR algorithm), the energy
counter, and the rocsparse_test driver with the reference's command line and
CROSSCHECK files written by gpu_test."""
import os
import subprocess
import sys

import numpy as np
import pytest

import gmr
from gmr import _lib
from helpers import rel_err, x_for
from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "g-mm-repair_amd", "bin")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _lib.vendor()


def test_rocsparse_products_match_the_oracle():
    from gmr import vendor
    m = gmr.Matrix.genotype(33, 3000, 2000, founders=32, classes=4, seg_mean=40.0)
    seq, V = m.arrays()
    A = vendor.Csr.from_matrix(m)
    assert A.nnz == m.nnz() and A.device_bytes() >= 8 * A.nnz
    x = x_for(0, 2000, 3)
    y_star, den = O.brute_force(0, 3000, 2000, seq, V, x)
    X = x_for(0, 2000, 4, B=8)
    Y_star, denB = O.brute_force_batched(0, 3000, 2000, seq, V, X)
    ran = 0
    for alg in vendor.ALGS:
        for XX, YY, dd in ((x, y_star, den), (X, Y_star, denB)):
            _say("rocsparse", alg, XX.shape)
            Xt = torch.from_numpy(XX).cuda()
            try:
                y = A.run_torch(Xt, alg=alg)
            except gmr.GmrError as e:
                # nnz_split is SpMM only; the default and the row algorithms must run
                assert alg == "nnz_split" and XX.ndim == 1, e
                continue
            torch.cuda.synchronize()
            assert rel_err(y.cpu().numpy(), YY, dd) <= 1e-5, (alg, XX.shape)
            # repeated calls (the analysis is cached) recompute into a cleared y
            y.zero_()
            A.run_torch(Xt, y, alg=alg)
            A.run_torch(Xt, y, alg=alg)
            torch.cuda.synchronize()
            assert rel_err(y.cpu().numpy(), YY, dd) <= 1e-5, (alg, XX.shape, "repeat")
            ran += 1
    assert ran == 7


def test_energy_counter():
    from gmr import vendor
    e0 = vendor.energy_uj(0)
    if e0 is None:
        pytest.skip("the energy counter cannot be read on this machine")
    A = vendor.Csr.from_matrix(gmr.Matrix.haplotype(3, 2048, 1024))
    mj, w = A.energy(B=4, seconds=1.0)  # the counter advances about once a second
    assert mj > 0 and 10 < w < 5000, (mj, w)
    assert vendor.energy_uj(0) > e0


def test_rocsparse_test_driver(tmp_path):
    m = gmr.Matrix.haplotype(78, 2500, 900, founders=24, seg_mean=48.0)
    g = m.repair()
    base = str(tmp_path / "g")
    g.write(base)
    prefix = str(tmp_path / "out")
    for sr, name in ((0, "plustimes"), (1, "boolean"), (2, "tropical")):
        x = x_for(sr, 900, 600 + sr)
        xf = tmp_path / f"x{sr}.bin"
        x.tofile(xf)
        env = dict(os.environ, SEMIRING=name, XVEC=str(xf), CROSSCHECK=prefix, GMR_ENERGY_SECONDS="0")
        _say("gpu_test dump", name)
        r = subprocess.run([os.path.join(BIN, "gpu_test"), base, "2500", "900", "1"], capture_output=True, text=True,
                           env=env, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        _say("rocsparse_test crosscheck", name)
        r = subprocess.run([os.path.join(BIN, "rocsparse_test"), base, "2500", "900", "1"], capture_output=True,
                           text=True, env=env, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        lines = [ln for ln in r.stdout.splitlines() if ln.startswith("CROSSCHECK") and "vs engine" in ln]
        assert lines and all("SUCCESS" in ln for ln in lines), r.stdout
    _say("rocsparse_test timing")
    env = dict(os.environ, SEMIRING="plustimes")
    env.pop("CROSSCHECK", None)
    r = subprocess.run([os.path.join(BIN, "rocsparse_test"), base, "2500", "900", "3", "repair", "4"],
                       capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = r.stdout
    assert f"NNZ: {m.nnz()}" in out and "SUCCESS: Verification passed!" in out
    assert "rocSPARSE SpMV completed. Average time:" in out and "ENERGY mJ/vector:" in out
    spmm = [ln for ln in out.splitlines() if ln.startswith("rocSPARSE SpMM (B=4")]
    assert spmm and all("SUCCESS" in ln for ln in spmm if "UNSUPPORTED" not in ln), out
    assert any("SUCCESS" in ln for ln in spmm)
    xf = prefix + ".plustimes.x"
    assert np.fromfile(xf, np.float32).shape == (900,)
