"""The learned proposer end to end on small grammars: timings from the sweep,
training on the GPU by the engine's own kernels, held-out choices scored
against exhaustive tuning (regret) and against the plain schedule (gain)."""
import json
import os

import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_proposer_trains_and_beats_the_plain_schedule(tmp_path):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gmr import train_proposer
    train_proposer.main(["--out", str(tmp_path), "--quick", "--epochs", "1500", "--reps", "2"])
    with open(os.path.join(tmp_path, "proposer_report.json")) as f:
        rep = json.load(f)
    tr, ho = rep["training"], rep["heldout"]
    print(json.dumps({"training": tr, "heldout": ho, "fused_ms": rep["fused_ms"]}))
    assert tr["loss_last"] < 0.2 * tr["loss_first"]
    assert ho["n"] == 6
    # What the network gains over the method without it: the plain schedule.
    assert ho["gain_over_plain_of_totals"] > 1.0
    assert all(r["regret"] >= -1e-9 for r in rep["per_grammar"])
