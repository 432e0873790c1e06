"""Launch-shape autotuning for the multi-RHS PCG.

The right-hand-side column tile decides the lanes per CSR row (4 columns per
lane, up to a wave per row), the work-group count, and how much of the
iteration's working set stays in the last-level cache. The best width depends
on the operator's size and on the number of right-hand sides. The tuner runs a
short, iteration-capped probe solve per candidate width on the real system
and right-hand sides, and keeps the width with the lowest device time per
column-iteration."""
import numpy as np

from . import hip


def tune_rhs_tile(system, B, k, trans=False, seg=None, candidates=(16, 32, 64, 128, 256), probe_iters=64,
                  repeats=2):
    """B: device block (N x k). Returns (best tile, 0 for one whole-width solve;
    {width: ms per column-iteration})."""
    ctx = system.ctx
    X = ctx.empty((system.ndof, k))
    cost = {}
    # widths at or above k all mean one whole-width solve: probe it once
    cands = sorted({min(t, k) for t in candidates})
    for t in cands:
        kt = t
        best = np.inf
        for _ in range(repeats):
            o = hip.default_opts(rhs_tile=0, max_inner=probe_iters, check_every=probe_iters)
            st = hip.solve_probe(system, _Cols(B, kt, k), _Cols(X, kt, k), kt, trans=trans, opts=o, seg=seg) \
                if kt < k else hip.solve_probe(system, B, X, k, trans=trans, opts=o, seg=seg)
            if st["inner_iters"] > 0:
                best = min(best, st["ms"] / (st["inner_iters"] * kt))
        cost[t] = best
    tile = min(cost, key=cost.get)
    return (0 if tile >= k else tile), cost


class _Cols:
    """The first kt columns of a row-major N x k device block, copied to a
    contiguous N x kt block (the probe solves a tile on its own)."""

    def __init__(self, A, kt, k):
        N = A.shape[0]
        h = A.download().reshape(N, k)[:, :kt]
        self._a = A.ctx.to_device(np.ascontiguousarray(h))
        self.ptr = self._a.ptr
