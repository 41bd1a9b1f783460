"""Row-block grammars (mm-repair's `-b` blocks) swept in one batch launch:
the product equals the oracle's brute force from the uncompressed matrix."""
import sys

import numpy as np
import pytest

import gmr
from gmr import _lib
from gmr.blocks import BlockGrammar
from helpers import rel_err, x_for
from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _lib.dev()


@pytest.mark.parametrize("nblocks,inline", [(1, 0), (3, 0), (7, 2)])
def test_block_grammars_match_the_oracle(nblocks, inline):
    m = gmr.Matrix.genotype(44, 3000, 2000, founders=32, classes=4, seg_mean=40.0)
    seq, V = m.arrays()
    bg = BlockGrammar(m, nblocks, inline_top=inline)
    assert bg.info()["blocks"] == nblocks
    for sr in (0, 1, 2):
        for B in (1, 16):
            X = x_for(sr, 2000, 10 * sr + B, B=None if B == 1 else B)
            Y_star, den = (O.brute_force if B == 1 else O.brute_force_batched)(sr, 3000, 2000, seq, V, X)
            for var in ("graph", "persistent", "levels"):
                print("blocks", nblocks, inline, sr, B, var, flush=True)
                sys.stdout.flush()
                Y = bg.run_torch(sr, torch.from_numpy(X).cuda(), variant=gmr.VARIANTS[var])
                torch.cuda.synchronize()
                if var == "persistent":
                    bg.batch.check()
                Y = Y.cpu().numpy()
                assert Y.shape == Y_star.shape
                if sr == 0:
                    assert rel_err(Y, Y_star, den) <= 1e-5
                else:
                    assert np.array_equal(Y.astype(np.int64), Y_star.astype(np.int64))
