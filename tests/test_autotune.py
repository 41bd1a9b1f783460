"""The exhaustive autotuner's search space and disk cache (CPU; the device
timing is replaced by a deterministic stand-in)."""
import gmr
from gmr import autotune


class _Engine:
    def info(self):
        return {"rows": 10, "cols": 4, "U": 5, "max_depth": 3}

    def set_gather_order(self, order):
        pass


class _X:
    def __init__(self, B):
        self.B = B

    def dim(self):
        return 1 if self.B == 1 else 2

    @property
    def shape(self):
        return (4, self.B)


def test_candidates_cover_variants_emits_tiles_and_orders():
    c64 = autotune.candidates(64, gmr.RESIDENT)
    descr = {autotune.describe(c) for c in c64}
    assert "levels/gather/tile=B/blk=256" in descr
    assert "graph/gather/tile=16/blk=256/order=rule" in descr
    assert "persistent/spot/tile=32/blk=1024" in descr
    assert len(descr) == len(c64)
    # one vector: no tiles below B, and the gather order does not apply
    c1 = autotune.candidates(1, gmr.RESIDENT)
    assert all(c["b_tile"] == 0 and c["gather_order"] == 0 for c in c1)
    # double-buffered layouts have no row gather
    assert all(c["emit"] == gmr.EMIT_SPOT for c in autotune.candidates(16, gmr.WAVE64))


def test_tune_picks_the_fastest_and_reuses_the_disk_cache(tmp_path, monkeypatch):
    calls = []

    def fake_time(engine, semiring, X, Y, cand, reps=3, warm=1):
        calls.append(autotune.describe(cand))
        return 1.0 + 0.1 * cand["b_tile"] + (0.5 if cand["emit"] == gmr.EMIT_SPOT else 0.0) + 0.01 * cand["block"] / 256

    monkeypatch.setattr(autotune, "time_candidate", fake_time)
    cache = str(tmp_path / "tune.json")
    cands = autotune.candidates(64, gmr.RESIDENT)
    best, res = autotune.tune(_Engine(), 0, _X(64), None, cands, cache_path=cache)
    assert len(calls) == len(cands) and len(res) == len(cands)
    assert best["emit"] == gmr.EMIT_GATHER and best["b_tile"] == 0 and best["block"] == 256
    n = len(calls)
    best2, res2 = autotune.tune(_Engine(), 0, _X(64), None, cands, cache_path=cache)
    assert len(calls) == n  # every timing came from the cache
    assert best2 == best and [ms for _, ms in res2] == [ms for _, ms in res]
    # another batch width is another key: timed again
    autotune.tune(_Engine(), 0, _X(16), None, autotune.candidates(16, gmr.RESIDENT), cache_path=cache)
    assert len(calls) > n


class _Engine2(_Engine):
    def info(self):
        return {"rows": 10, "cols": 4, "U": 5, "max_depth": 2}


def test_tune_forms_keeps_the_fastest_form(monkeypatch):
    def fake_time(engine, semiring, X, Y, cand, reps=3, warm=1):
        return (0.5 if isinstance(engine, _Engine2) else 1.0) + 0.1 * cand["b_tile"]

    monkeypatch.setattr(autotune, "time_candidate", fake_time)
    forms = [{"label": "as built", "engine": _Engine()}, {"label": "inlined", "engine": _Engine2()}]
    cands = autotune.candidates(64, gmr.RESIDENT)
    form, best, tuned = autotune.tune_forms(forms, 0, _X(64), None, cands)
    assert form["label"] == "inlined" and best["b_tile"] == 0
    assert len(tuned) == 2 * len(cands) and {lab for lab, _, _ in tuned} == {"as built", "inlined"}
    assert min(ms for _, _, ms in tuned) == 0.5



def test_candidates_for_narrow_tiles_cover_both_gathers():
    """Tiles of 2-4 vectors: the XCD-partitioned gather (the default, with
    its pass counts) and the row segments in both orders."""
    for B in (2, 3, 4):
        d = {autotune.describe(c) for c in autotune.candidates(B, gmr.RESIDENT)}
        for p in ("auto", "1", "2", "4"):
            assert f"levels/gather/tile=B/blk=256/xcd-passes={p}" in d
        assert "levels/gather/tile=B/blk=256/segments" in d
        assert "levels/gather/tile=B/blk=256/order=rule/segments" in d
    d16 = {autotune.describe(c) for c in autotune.candidates(16, gmr.RESIDENT)}
    assert not any("xcd-passes" in x or "segments" in x for x in d16)


def test_bit_mode_candidates_only_for_boolean_single_vectors():
    """The Boolean bit mode is searched for one vector with gather emission in
    the levels and graph variants, with and without paired levels, and
    nowhere else."""
    d = {autotune.describe(c) for c in autotune.candidates(1, gmr.RESIDENT, semiring=gmr.BOOLEAN)}
    bits = {x for x in d if "bits" in x}
    for v in ("levels", "graph"):
        for r in (262144, 524288, 1048576):
            for ch in (4, 8):
                assert f"{v}/gather/tile=B/blk=256/bits(range={r})/chunks={ch}" in bits
                assert f"{v}/gather/tile=B/blk=256/bits(range={r})/paired-levels/chunks={ch}" in bits
    assert len(bits) == 24
    for B, sr, lay in ((1, gmr.PLUS_TIMES, gmr.RESIDENT), (1, gmr.TROPICAL, gmr.RESIDENT), (4, gmr.BOOLEAN, gmr.RESIDENT),
                       (1, gmr.BOOLEAN, gmr.WAVE64), (1, None, gmr.RESIDENT)):
        assert not any(c.get("flags", 0) & gmr.RUN_BITS for c in autotune.candidates(B, lay, semiring=sr))


def test_candidates_search_the_level_chunk_sizes():
    """One vector: the XCD-partitioned gather's chunks per wave, the LDS-range
    gather's range sizes and chunks with and without the row order (levels and
    graph, every semiring), and
    the LDS-staged node update's windows; several vectors: none of these."""
    for sr in (gmr.PLUS_TIMES, gmr.BOOLEAN, gmr.TROPICAL):
        d = {autotune.describe(c) for c in autotune.candidates(1, gmr.RESIDENT, semiring=sr)}
        for v in ("levels", "graph"):
            for span in (16384, 32768, 40960):
                for ch in (2, 4):
                    assert f"{v}/gather/tile=B/blk=256/lds-ranges(values={span})/chunks={ch}" in d
                    assert f"{v}/gather/tile=B/blk=256/lds-ranges(values={span})/row-order/chunks={ch}" in d
            for win in (2048, 4096, 16384):
                assert f"{v}/spot/tile=B/blk=256/eval=lds(window={win})" in d
        for ch in (2, 4):
            assert f"levels/gather/tile=B/blk=256/xcd-passes=2/chunks={ch}" in d
            assert f"graph/gather/tile=B/blk=256/lds-ranges(values=40960)/row-order/fused-tail/chunks={ch}" in d
        assert "levels/spot/tile=B/blk=256/fused-tail" in d
        assert "levels/gather/tile=B/blk=256/lds-ranges(values=40960)/row-order/fused-tail/paired-levels/chunks=4" in d
        assert not any("paired-levels" in x for x in d if "/spot/" in x)
        bits = {x for x in d if "/bits(" in x}
        assert bool(bits) == (sr == gmr.BOOLEAN)
        assert all(x.replace("/paired-levels", "") in d for x in bits)
        if sr == gmr.BOOLEAN:
            assert "levels/gather/tile=B/blk=256/bits(range=262144)/paired-levels/chunks=8" in d
        assert not any("fused-tail" in x for x in d if x.startswith("persistent") or "/bits(" in x)
        assert not any("lds-ranges" in x for x in d if x.startswith("persistent"))
    for B in (4, 64):
        d = {autotune.describe(c) for c in autotune.candidates(B, gmr.RESIDENT)}
        assert not any("lds" in x or "chunks" in x or "fused-tail" in x for x in d)
    assert not any("fused-tail" in autotune.describe(c) for c in autotune.candidates(1, gmr.WAVE64))
