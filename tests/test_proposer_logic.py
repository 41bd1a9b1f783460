"""Decision logic of the schedule proposer on synthetic timing tables (CPU)."""
import math

import numpy as np

from gmr import proposer as P


def _table():
    # two tiles, 3 levels; plan 3 (graph/gather) is fastest at tile 64
    ev64 = np.array([[1.0, 0.9, 3.0, np.nan], [2.0, 2.1, 5.0, np.nan], [0.5, 0.4, 0.6, np.nan]])
    ev16 = ev64 * 0.3
    return {64: {"eval": ev64, "plan": np.array([10.0, 6.0, 9.0, 5.0, 8.0, 5.5])},
            16: {"eval": ev16, "plan": np.array([3.0, 1.9, 2.8, 1.7, 2.5, 1.8])}}


def test_sweep_cost_counts_passes_and_kernel_deltas():
    t = _table()
    c = P.sweep_cost(t, 64, 16, 3, [1, 0, 1])
    delta = (0.9 - 1.0) * 0.3 + 0.0 + (0.4 - 0.5) * 0.3
    assert math.isclose(c, 4 * (1.7 + delta))


def test_exhaustive_and_heuristic():
    t = _table()
    cost, Bt, plan, kinds = P.exhaustive(t, 64)
    assert (Bt, plan, kinds) == (64, 3, [1, 0, 1])
    assert math.isclose(cost, 5.0 - 0.1 - 0.1)
    h = P.heuristic(t, 64)
    assert h[1] == 64 and h[2] == P.DEFAULT_PLAN and h[3] == [P.EVAL_THREAD] * 3
    assert math.isclose(h[0], 10.0 + (0.9 - 1.0) + (2.1 - 2.0) + (0.4 - 0.5))


def test_regret_of_the_optimum_is_zero():
    t = _table()
    c, Bt, plan, kinds = P.exhaustive(t, 64)
    assert P.regret(t, 64, (Bt, plan, kinds, 0.0)) == 0.0
    assert P.regret(t, 64, (64, 0, kinds, 0.0)) > 0.0


def test_features_shape_and_tiles():
    assert P.tile_candidates(100) == [16, 32, 64, 100]
    assert P.tile_candidates(1) == [1]
    import gmr
    m = gmr.Matrix.haplotype(3, 300, 128, founders=8, seg_mean=16.0)
    g = m.repair()
    arr = g.schedule(300, 128, "resident").arrays()
    F, gs = P.features(arr, 16)
    assert F.shape == (arr["info"]["max_depth"], P.N_FEAT) and gs.shape == (P.N_FEAT,)
    assert np.all(np.isfinite(F)) and np.all(np.isfinite(gs))
    assert gs[0] == 1.0 and np.all(F[:, 0] == 0.0)


def test_shipped_weights_match_the_network():
    """The trained networks shipped with the package have one parameter per
    weight and bias of the 12 -> 64 -> 64 -> 10 MLP and finite standardisation."""
    for path in (P.DEFAULT_WEIGHTS, P.CFG123_WEIGHTS):
        with np.load(path) as f:
            p, mu, sd = f["params"], f["mu"], f["sd"]
            tsd = float(f["tsd"])
        H, I, O = P.HIDDEN, P.N_FEAT, P.N_OUT
        assert p.shape == (I * H + H + H * H + H + H * O + O,)
        assert mu.shape == sd.shape == (I,)
        assert np.isfinite(p).all() and (sd > 0).all() and tsd > 0


def test_run_options_from_a_choice():
    Bt, variant, opts = P.run_options((32, 5, [1, 0, 2], 0.1))
    assert Bt == 32 and variant == P.PLANS[5][0] and opts["emit"] == P.PLANS[5][1]
    assert opts["level_kind"].tolist() == [0, 1, 0, 2] and opts["b_tile"] == 32


def _syn_table(tiles, L=3, plan_ms=None, eval_ms=None):
    """A synthetic timing table: per tile, [L, 4] level times and 6 plan times."""
    t = {}
    for Bt in tiles:
        ev = np.array(eval_ms if eval_ms is not None else [[1.0, 2.0, 3.0, 0.5]] * L, dtype=float)
        if Bt != 1:
            ev[:, P.EVAL_LDS] = np.nan
        t[Bt] = {"eval": ev, "plan": np.array(plan_ms if plan_ms is not None else [6, 5, 4, 3, 2, 1.0]) * Bt}
    return t


def test_constant_rules_cost_and_best():
    tab = _syn_table([1, 16, 40])
    # plan 5 (persistent/gather), group kernel, tile = B: 1 pass of the B-wide plan time
    assert P.constant_cost(tab, 40, (5, P.EVAL_GROUP, 0)) == 40.0
    # tile min(B, 16) = 16: three passes of the 16-wide sweep
    assert P.constant_cost(tab, 40, (5, P.EVAL_GROUP, 16)) == 3 * 16.0
    # a tile the table has no timing for: not applicable
    assert P.constant_cost(tab, 40, (5, P.EVAL_GROUP, 32)) is None
    # the LDS kernel only exists for one vector; elsewhere the group kernel stands in
    assert P.constant_cost(tab, 40, (5, P.EVAL_LDS, 0)) == P.constant_cost(tab, 40, (5, P.EVAL_GROUP, 0))
    assert P.constant_cost(tab, 1, (5, P.EVAL_LDS, 0)) == 1.0 + 3 * (0.5 - 1.0)
    # the best rule over several problems only uses rules every problem has timings for
    rule, tot = P.best_constant([(tab, 40), (tab, 1)])
    assert rule[0] == 5 and rule[1] == P.EVAL_LDS and rule[2] in (0, 16, 64, 128)
    assert abs(tot - (40.0 + (1.0 - 1.5))) < 1e-12
    assert P.describe_rule((5, P.EVAL_LDS, 16)) == "persistent/gather/lds/tile=min(B,16)"
    assert len(P.constant_rules()) == len(P.PLANS) * 4 * len(P.CONST_TILES)
