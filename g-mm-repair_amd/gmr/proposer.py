"""Learned schedule proposer.

For a grammar and a batch width B it picks, without timing anything, a launch
plan (per-level launches, one hipGraph, or one persistent launch; emission on
the spot or by row gather), the node-update kernel of every level (group /
thread / wave / LDS-staged) and the vector tile Bt. One small MLP
(12 features -> 64 -> 64 -> 10, ReLU) does both: for a level sample it predicts
the log device time of each of the four node-update kernels, for a grammar
sample the log time of a whole sweep under each of the six plans. The choice
minimises ceil(B/Bt) * (predicted plan time, corrected by the chosen kernels'
predicted gain per level). The network is trained on the GPU by the engine's own
kernels (gmr_mlp_*) on timings the sweep itself produces. Exhaustive timing of
every (plan, level, kernel, tile) stays available as the fallback and as the
yardstick for the proposer's regret. Every choice computes the same Y.
"""
import ctypes as C
import math
import os

import numpy as np

from . import (EMIT_GATHER, EMIT_SPOT, EVAL_GROUP, EVAL_LDS, EVAL_THREAD, EVAL_WAVE, VARIANT_GRAPH,
               VARIANT_LEVELS, VARIANT_PERSISTENT, GmrError, _check, _lib)

KINDS = (EVAL_GROUP, EVAL_THREAD, EVAL_WAVE, EVAL_LDS)
PLANS = [(v, e) for v in (VARIANT_LEVELS, VARIANT_GRAPH, VARIANT_PERSISTENT) for e in (EMIT_SPOT, EMIT_GATHER)]
PLAN_NAMES = [f"{['levels', 'graph', 'persistent'][v]}/{['spot', 'gather'][e]}" for v, e in PLANS]
DEFAULT_PLAN = 0  # levels / spot: the plain per-level sweep
N_FEAT = 12
N_OUT = 4 + len(PLANS)
HIDDEN = 64


def tile_candidates(B):
    """Vector tiles considered for B right-hand sides."""
    return sorted({B} | {t for t in (16, 32, 64, 128) if t < B})


def _log2p(x):
    return math.log2(1.0 + float(x))


def grammar_stats(arr):
    """Per-slot reference counts (rule children and top-sequence uses) of a
    Resident-layout schedule, used by the popularity feature."""
    NL = len(arr["left"])
    pop = np.zeros(max(NL, 1), np.int64)
    for c in (arr["left"], arr["right"]):
        nt = c[c >= 0]
        if len(nt):
            pop += np.bincount(nt, minlength=len(pop))[: len(pop)]
    if len(arr["emit_pos"]):
        pop += np.bincount(arr["emit_pos"], minlength=len(pop))[: len(pop)]
    return np.log2(1.0 + pop)


def features(arr, Bt, lpop=None):
    """(F_eval [L, N_FEAT], f_gather [N_FEAT]) for a Resident-layout schedule
    (arrays from Schedule.arrays()) swept with tile width Bt."""
    if lpop is None:
        lpop = grammar_stats(arr)
    lo_ = arr["level_offsets"]
    eo = arr["emit_offsets"]
    L = len(lo_) - 1
    vec = 1.0 if Bt % 4 == 0 else 0.0
    F = np.zeros((L, N_FEAT), np.float32)
    for k in range(1, L + 1):
        lo, hi = int(lo_[k - 1]), int(lo_[k])
        w = hi - lo
        lft, rgt = arr["left"][lo:hi], arr["right"][lo:hi]
        ntl = lft[lft >= 0]
        nt = np.concatenate([ntl, rgt[rgt >= 0]])
        span = (int(ntl.max()) - int(ntl.min()) + 1) if len(ntl) else 0
        gap = float(np.abs(np.diff(ntl)).mean()) if len(ntl) > 1 else 0.0
        pop = float(lpop[nt].mean()) if len(nt) else 0.0
        ne = int(eo[k] - eo[k - 1])
        F[k - 1] = [0.0, _log2p(w), math.log2(Bt), vec, float((lft < 0).mean()) if w else 0.0,
                    float((rgt < 0).mean()) if w else 0.0, _log2p(span / max(w, 1)), _log2p(gap), _log2p(lo),
                    pop, _log2p(w * Bt), _log2p(ne)]
    info = arr["info"]
    rows = info["rows"]
    E, Et = info["emit_total"], info["num_term_emit"]
    ep = arr["emit_pos"]
    g = np.array([1.0, _log2p(rows), math.log2(Bt), vec, Et / max(E + Et, 1), _log2p(len(arr["left"])),
                  _log2p((E + Et) / max(rows, 1)), L / 32.0, _log2p(info["max_width"]),
                  float(lpop[ep].mean()) if len(ep) else 0.0, _log2p((E + Et) * Bt),
                  _log2p(len(arr["left"]) * Bt)], np.float32)
    return F, g


def _time_plan(engine, X, Y, variant, emit, reps):
    import torch
    kw = dict(variant=variant, emit=emit)
    engine.run_torch(0, X, Y, **kw)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        engine.run_torch(0, X, Y, **kw)
    e1.record()
    torch.cuda.synchronize()
    if variant == VARIANT_PERSISTENT:
        engine.check()
    return e0.elapsed_time(e1) / reps


def collect(engine, arr, tiles, reps=3, seed=0):
    """Time, for each tile width Bt (Resident layout): every level under every
    node-update kernel, and a whole sweep under every plan (default kernels).
    Returns a table {Bt: {"eval": [L, 4] ms (nan where a kernel does not
    apply), "plan": [6] ms}} and training samples (X, T, M)."""
    import torch
    info = engine.info()
    rows, cols = info["rows"], info["cols"]
    lpop = grammar_stats(arr)
    gen = torch.Generator(device="cuda").manual_seed(seed)
    table, Xs, Ts, Ms = {}, [], [], []
    L = info["max_depth"]
    for Bt in tiles:
        X = torch.randn(cols, Bt, device="cuda", generator=gen) if Bt > 1 else torch.randn(cols, device="cuda",
                                                                                                generator=gen)
        Y = torch.empty((rows, Bt) if Bt > 1 else (rows,), device="cuda")
        ev = np.full((L, 4), np.nan)
        for kind in KINDS:
            if kind == EVAL_LDS and Bt != 1:
                continue
            t = engine.time_levels(0, X, Y, reps=reps, eval_kind=kind, emit=EMIT_GATHER)
            ev[:, kind] = t[1:L + 1]
        plan = np.array([_time_plan(engine, X, Y, v, e, reps) for v, e in PLANS])
        table[Bt] = {"eval": ev, "plan": plan}
        F, g = features(arr, Bt, lpop)
        T = np.zeros((L + 1, N_OUT), np.float32)
        M = np.zeros((L + 1, N_OUT), np.float32)
        T[:L, :4] = np.log(np.maximum(np.nan_to_num(ev, nan=1.0), 1e-6) * 1e3)
        M[:L, :4] = np.isfinite(ev)
        T[L, 4:] = np.log(np.maximum(plan, 1e-6) * 1e3)
        M[L, 4:] = 1.0
        Xs.append(np.vstack([F, g[None, :]]))
        Ts.append(T)
        Ms.append(M)
    return table, (np.vstack(Xs), np.vstack(Ts), np.vstack(Ms))


def sweep_cost(table, B, Bt, plan, kinds):
    """Measured cost (ms) of B vectors with tile Bt, a plan and per-level
    kernels: the plan's sweep time (default kernels) plus each level's measured
    difference between its chosen kernel and the default one, per tile pass."""
    t = table[Bt]
    ev = t["eval"]
    delta = sum(float(ev[k, kinds[k]] - ev[k, EVAL_GROUP]) for k in range(ev.shape[0]))
    return math.ceil(B / Bt) * (float(t["plan"][plan]) + delta)


def exhaustive(table, B):
    """Best (cost, Bt, plan, kinds) by measured times: the autotuning fallback."""
    best = None
    for Bt in table:
        ev = table[Bt]["eval"]
        kinds = [int(np.nanargmin(ev[k])) for k in range(ev.shape[0])]
        for plan in range(len(PLANS)):
            c = sweep_cost(table, B, Bt, plan, kinds)
            if best is None or c < best[0]:
                best = (c, Bt, plan, kinds)
    return best


def heuristic(table, B):
    """The fixed method without the network: the plain sweep (per-level
    launches, emission on the spot), thread-per-rule everywhere, tile = B."""
    kinds = [EVAL_THREAD] * table[B]["eval"].shape[0]
    return sweep_cost(table, B, B, DEFAULT_PLAN, kinds), B, DEFAULT_PLAN, kinds


# Constant schedules: one plan, one node-update kernel on every level and one
# tile rule (0: tile = B; T: tile = min(B, T)) for every grammar and B -- the
# best of these is the method without the network to beat.
CONST_TILES = (0, 16, 32, 64, 128)


def constant_rules():
    return [(plan, kind, T) for plan in range(len(PLANS)) for kind in KINDS for T in CONST_TILES]


def constant_cost(table, B, rule):
    """Measured cost (ms) of a constant rule for B vectors, or None when the
    table has no timing of the tile it asks for. A kernel that does not apply
    to a level (LDS staging for tiles wider than one vector) falls back to the
    group kernel there, as the engine does."""
    plan, kind, T = rule
    Bt = B if T == 0 else min(B, T)
    if Bt not in table:
        return None
    ev = table[Bt]["eval"]
    kinds = [kind if np.isfinite(ev[k, kind]) else EVAL_GROUP for k in range(ev.shape[0])]
    return sweep_cost(table, B, Bt, plan, kinds)


def best_constant(problems):
    """The constant rule with the lowest summed cost over problems [(table, B)]
    among the rules every problem has timings for: (rule, total ms)."""
    best = None
    for rule in constant_rules():
        tot = 0.0
        for table, B in problems:
            c = constant_cost(table, B, rule)
            if c is None:
                tot = None
                break
            tot += c
        if tot is not None and (best is None or tot < best[1]):
            best = (rule, tot)
    return best


def describe_rule(rule):
    plan, kind, T = rule
    return f"{PLAN_NAMES[plan]}/{['group', 'thread', 'wave', 'lds'][kind]}/tile={'B' if T == 0 else f'min(B,{T})'}"


class Mlp:
    """The network on the GPU (gmr_mlp_*); inputs are standardised features."""

    def __init__(self, nin=N_FEAT, nh=HIDDEN, nout=N_OUT, seed=1):
        h = C.c_void_p()
        _check(_lib.dev().gmr_mlp_create(nin, nh, nout, seed, C.byref(h)))
        self._h, self.nin, self.nh, self.nout = h, nin, nh, nout

    @property
    def nparams(self):
        return int(_lib.dev().gmr_mlp_num_params(self._h))

    def get_params(self):
        p = np.zeros(self.nparams, np.float32)
        _check(_lib.dev().gmr_mlp_get_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float))))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert len(p) == self.nparams
        _check(_lib.dev().gmr_mlp_set_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float))))

    @staticmethod
    def _dev(a):
        import torch
        return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()

    def forward(self, X):
        import torch
        Xd = self._dev(X)
        out = torch.empty((X.shape[0], self.nout), device="cuda")
        _check(_lib.dev().gmr_mlp_forward(self._h, C.c_void_p(Xd.data_ptr()), X.shape[0],
                                          C.c_void_p(out.data_ptr()), None))
        torch.cuda.synchronize()
        return out.cpu().numpy()

    def loss_grad(self, X, T, M):
        Xd, Td, Md = self._dev(X), self._dev(T), self._dev(M)
        loss = C.c_float()
        g = np.zeros(self.nparams, np.float32)
        _check(_lib.dev().gmr_mlp_loss_grad(self._h, C.c_void_p(Xd.data_ptr()), C.c_void_p(Td.data_ptr()),
                                            C.c_void_p(Md.data_ptr()), X.shape[0], float(M.sum()), C.byref(loss),
                                            g.ctypes.data_as(C.POINTER(C.c_float))))
        return loss.value, g

    def train(self, X, T, M, epochs=2000, lr=3e-3):
        Xd, Td, Md = self._dev(X), self._dev(T), self._dev(M)
        losses = np.zeros(epochs, np.float32)
        _check(_lib.dev().gmr_mlp_train(self._h, C.c_void_p(Xd.data_ptr()), C.c_void_p(Td.data_ptr()),
                                        C.c_void_p(Md.data_ptr()), X.shape[0], float(M.sum()), epochs, lr,
                                        losses.ctypes.data_as(C.POINTER(C.c_float))))
        return losses

    def __del__(self):
        lib = getattr(_lib, "_dev", None) if _lib is not None else None
        if getattr(self, "_h", None) and lib is not None:
            lib.gmr_mlp_destroy(self._h)
            self._h = None


class Proposer:
    """Standardisation + network + decision rule."""

    def __init__(self, seed=1):
        self.net = Mlp(seed=seed)
        self.mu = np.zeros(N_FEAT, np.float32)
        self.sd = np.ones(N_FEAT, np.float32)
        self.tmu = 0.0
        self.tsd = 1.0
        # The constant schedule the network has to beat (the best one on the
        # training timings) and the predicted gain it needs to be followed.
        self.baseline = None
        self.margin = 0.0

    def fit(self, X, T, M, epochs=3000, lr=3e-3):
        self.mu = X.mean(0).astype(np.float32)
        self.sd = (X.std(0) + 1e-3).astype(np.float32)
        vals = T[M > 0]
        self.tmu = float(vals.mean())
        self.tsd = float(vals.std() + 1e-6)
        Tn = np.where(M > 0, (T - self.tmu) / self.tsd, 0.0).astype(np.float32)
        return self.net.train((X - self.mu) / self.sd, Tn, M, epochs=epochs, lr=lr)

    def predict(self, X):
        """Predicted times in ms, [N, N_OUT]."""
        o = self.net.forward((X - self.mu) / self.sd)
        return np.exp(o * self.tsd + self.tmu) / 1e3

    def _predicted(self, arr, Bt, lpop, cache):
        """Predicted times [L + 1, N_OUT] for tile Bt (cached per tile: they do
        not depend on B)."""
        if cache is not None and Bt in cache:
            return cache[Bt]
        F, g = features(arr, Bt, lpop)
        P = self.predict(np.vstack([F, g[None, :]]))
        if cache is not None:
            cache[Bt] = P
        return P

    def _predict_rule(self, arr, B, rule, lpop, cache=None):
        """(Bt, plan, kinds, predicted ms) of a constant rule."""
        plan, kind, T = rule
        Bt = B if T == 0 else min(B, T)
        P = self._predicted(arr, Bt, lpop, cache)
        k = EVAL_GROUP if (kind == EVAL_LDS and Bt != 1) else kind
        delta = float((P[:-1, k] - P[:-1, EVAL_GROUP]).sum())
        return Bt, plan, [k] * (P.shape[0] - 1), math.ceil(B / Bt) * (float(P[-1, 4 + plan]) + delta)

    def choose(self, arr, B, lpop=None, tiles=None, use_baseline=True, cache=None):
        """(Bt, plan, kinds per level, predicted ms) for B vectors. With a
        baseline rule set, the network's choice is taken only when it is
        predicted to beat the baseline's predicted cost by the margin;
        otherwise the baseline rule is the choice. `cache` (a dict) keeps the
        predictions of one grammar across calls."""
        if lpop is None:
            lpop = grammar_stats(arr)
        best = self._choose_net(arr, B, lpop, tiles, cache)
        if use_baseline and self.baseline is not None:
            base = self._predict_rule(arr, B, self.baseline, lpop, cache)
            if not best[3] < (1.0 - self.margin) * base[3]:
                return base
        return best

    def _choose_net(self, arr, B, lpop, tiles, cache=None):
        best = None
        for Bt in (tiles or tile_candidates(B)):
            P = self._predicted(arr, Bt, lpop, cache)
            ev = P[:-1, :4].copy()
            if Bt != 1:
                ev[:, EVAL_LDS] = np.inf
            kinds = [int(np.argmin(ev[k])) for k in range(ev.shape[0])]
            gain = float((ev.min(1) - ev[:, EVAL_GROUP]).sum())
            plan = int(np.argmin(P[-1, 4:]))
            t = math.ceil(B / Bt) * (float(P[-1, 4 + plan]) + gain)
            if best is None or t < best[3]:
                best = (Bt, plan, kinds, t)
        return best

    def state(self):
        st = {"params": self.net.get_params(), "mu": self.mu, "sd": self.sd,
              "tmu": np.float32(self.tmu), "tsd": np.float32(self.tsd)}
        if self.baseline is not None:
            st["baseline"] = np.array(self.baseline, np.int32)
            st["margin"] = np.float32(self.margin)
        return st

    def load_state(self, st):
        self.net.set_params(st["params"])
        self.mu, self.sd = st["mu"].astype(np.float32), st["sd"].astype(np.float32)
        self.tmu, self.tsd = float(st["tmu"]), float(st["tsd"])
        if "baseline" in st:
            self.baseline = tuple(int(v) for v in st["baseline"])
            self.margin = float(st["margin"])
        else:
            self.baseline, self.margin = None, 0.0

    def fit_margin(self, grammars, Bs, margins=(0.0, 0.05, 0.1, 0.2, 0.3, 0.5), log=None):
        """Pick the margin with the lowest summed measured cost over training
        grammars [(arr, table)] at every B in Bs (tiles limited to those the
        table timed). Returns {margin: total ms}."""
        totals = {m: 0.0 for m in margins}
        for i, (arr, table) in enumerate(grammars):
            lpop = grammar_stats(arr)
            cache = {}
            for m in margins:
                self.margin = m
                for B in Bs:
                    c = self.choose(arr, B, lpop=lpop, tiles=[t for t in table if t <= B], cache=cache)
                    totals[m] += sweep_cost(table, B, c[0], c[1], c[2])
            if log:
                log(f"margin fit: grammar {i + 1} of {len(grammars)}")
        self.margin = min(totals, key=totals.get)
        return totals

    def save(self, path):
        np.savez(path, **self.state())

    @classmethod
    def load(cls, path):
        p = cls()
        with np.load(path) as f:
            p.load_state({k: f[k] for k in f.files})
        return p

    @classmethod
    def default(cls):
        """The network shipped with the package: trained on MI355X sweep
        timings of haplotype and genotype grammars (including members generated
        with weak-compression parameters) and R-MAT graphs
        (gmr.train_proposer --vary --graphs), at B in {1, 3, 4, 16, 29, 64,
        100, 256}. CFG123_WEIGHTS holds the network trained on the haplotype
        and genotype families alone (--vary)."""
        return cls.load(DEFAULT_WEIGHTS)


DEFAULT_WEIGHTS = os.path.join(os.path.dirname(__file__), "proposer_mi355x.npz")
CFG123_WEIGHTS = os.path.join(os.path.dirname(__file__), "proposer_mi355x_cfg123.npz")


def run_options(choice):
    """(Bt, variant, opts) for Engine.run / Batch.run from a choose() result."""
    Bt, plan, kinds, _ = choice
    variant, emit = PLANS[plan]
    return Bt, variant, {"emit": emit, "b_tile": Bt, "level_kind": np.array([0] + list(kinds), np.int32)}


def regret(table, B, choice):
    """Measured cost of the proposer's choice over the exhaustive optimum, minus 1."""
    Bt, plan, kinds, _ = choice
    if Bt not in table:
        raise GmrError(f"tile {Bt} was not timed")
    best = exhaustive(table, B)[0]
    return sweep_cost(table, B, Bt, plan, kinds) / best - 1.0
