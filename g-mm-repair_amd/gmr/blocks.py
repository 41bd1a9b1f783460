"""Row-block grammars: a matrix compressed as n grammars over contiguous row
blocks -- what mm-repair's `matrepair -b n` builds (run_benchmark_16.py:88-91,
REPRODUCIBILITY.md:450) -- and swept together in one batch launch. The blocks'
rows are disjoint, so every grammar writes its own rows of one Y; the
product is the same matrix product as the single grammar's."""
import time

from . import EMIT_GATHER, VARIANT_GRAPH, Batch, SEMIRINGS, PLUS_TIMES


class BlockGrammar:
    def __init__(self, m, nblocks, inline_top=0, layout="resident"):
        if not 1 <= nblocks <= max(1, m.rows):
            raise ValueError(f"nblocks must be in [1, rows={m.rows}]")
        self.rows, self.cols = m.rows, m.cols
        self.layout = layout
        self.blocks = []  # (r0, r1, grammar, schedule, engine)
        t = time.time()
        for r0, r1, mb in m.row_blocks(nblocks):
            g = mb.repair()
            if inline_top:
                g = g.inline_top(inline_top)
            s = g.schedule(r1 - r0, m.cols, layout)
            self.blocks.append((r0, r1, g, s, s.engine()))
        self.build_seconds = time.time() - t
        self.batch = Batch([b[4] for b in self.blocks])

    def info(self):
        gi = [b[2].info() for b in self.blocks]
        return dict(blocks=len(self.blocks), rules=sum(i["nrules"] for i in gi), clen=sum(i["clen"] for i in gi),
                    depth=max(b[3].info()["max_depth"] for b in self.blocks))

    def run_torch(self, semiring, X, Y=None, variant=VARIANT_GRAPH, opts=None):
        """Y = M X for a float32 CUDA X of shape (cols,) or (cols, B); every
        block's grammar writes its rows of Y in one batch launch."""
        import torch
        if isinstance(semiring, str):
            semiring = SEMIRINGS[semiring]
        assert X.is_cuda and X.dtype == torch.float32 and X.is_contiguous() and X.shape[0] == self.cols
        X2 = X.reshape(self.cols, -1)
        B = X2.shape[1]
        dt = torch.float32 if semiring == PLUS_TIMES else torch.int32
        if Y is None:
            Y = torch.empty((self.rows, B) if X.dim() == 2 else (self.rows,), dtype=dt, device=X.device)
        Y2 = Y.reshape(self.rows, -1)
        Ys = [Y2[r0:r1] for r0, r1, *_ in self.blocks]
        if opts is None:
            opts = [{"emit": EMIT_GATHER} if self.layout == "resident" else {} for _ in self.blocks]
        self.batch.run_torch(semiring, [X2] * len(self.blocks), Ys, variant=variant, opts=opts)
        return Y
