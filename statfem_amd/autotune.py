"""Launch-shape autotuning for the multi-RHS PCG, with persisted choices.

Two launch-shape choices decide how a multi-RHS solve maps onto the GPU:

* lanes per node (`sfh_solve_opts.lanes`): each lane carries 4 right-hand
  sides, so a column tile is 4 x lanes columns. Wide tiles amortise the
  operator's coefficient loads over more columns; narrow tiles let the
  work-groups of a batch skip the padding columns of its narrower problems
  and the columns that have already converged.
* the RHS tile (`sfh_solve_opts.rhs_tile`): independent solves of that many
  columns one after another, whose working sets can stay in the last-level
  cache.

The best pair depends on the operator's size, the batch's segment layout and
the column count. The tuner runs one short, iteration-capped probe solve per
candidate on the real system and right-hand sides, keeps the candidate with
the lowest device time, and stores the choice in a JSON file keyed by the
device and the problem's shape, so later runs (and later processes) reuse it
without probing."""
import hashlib
import json
import os
import tempfile

import numpy as np

from . import hip

DEFAULT_LANES = (4, 8, 16)
DEFAULT_TILES = (0,)


def cache_path():
    """File of persisted choices: $STATFEM_AMD_TUNE_CACHE, else
    ~/.cache/statfem_amd/launch_shapes.json"""
    p = os.environ.get("STATFEM_AMD_TUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "statfem_amd", "launch_shapes.json")


def load_cache(path=None):
    path = path or cache_path()
    try:
        with open(path) as f:
            d = json.load(f)
        return d if isinstance(d, dict) else {}
    except (OSError, ValueError):
        return {}


def store_choice(key, choice, path=None):
    """Merge one choice into the cache file (atomic replace, so a concurrent
    reader sees the old or the new file, never a partial one)."""
    path = path or cache_path()
    d = load_cache(path)
    d[key] = choice
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    fd, tmp = tempfile.mkstemp(dir=os.path.dirname(os.path.abspath(path)), suffix=".tmp")
    with os.fdopen(fd, "w") as f:
        json.dump(d, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def problem_key(info, ndof, k, trans, seg=None, tol=None, extra=""):
    """Cache key: device (architecture, compute units), rows, columns,
    transposition, the batch's segment bounds (hashed) and the tolerance."""
    h = hashlib.sha1(np.ascontiguousarray(seg if seg is not None else [0, ndof], dtype=np.int64).tobytes())
    return (f"{info['arch']}/cu{info['num_cu']}/n{ndof}/k{k}/t{int(bool(trans))}/seg{h.hexdigest()[:12]}"
            f"/tol{tol if tol is not None else 0:g}{'/' + extra if extra else ''}")


def tune_solve(system, B, k, trans=False, seg=None, lanes=DEFAULT_LANES, tiles=DEFAULT_TILES, probe_iters=64,
               repeats=2, tol=1e-10, use_cache=True, path=None):
    """Best (lanes, rhs_tile) for solving `system` with the N x k device block B.

    Returns (choice, cost): choice = {"lanes": L, "rhs_tile": T, "source":
    "cache" | "probe"}; cost maps "L/T" to the best device ms of the capped
    probe solve (empty when the choice came from the cache)."""
    ctx = system.ctx
    key = problem_key(ctx.info(), system.ndof, k, trans, seg, tol)
    if use_cache:
        hit = load_cache(path).get(key)
        if hit and "lanes" in hit and "rhs_tile" in hit:
            return {"lanes": int(hit["lanes"]), "rhs_tile": int(hit["rhs_tile"]), "source": "cache"}, {}
    X = ctx.empty((system.ndof, k))
    cost = {}
    # tiles at or above k all mean one whole-width solve: probe it once
    cand_tiles = sorted({0 if (t <= 0 or t >= k) else int(t) for t in tiles})
    for t in cand_tiles:
        for L in lanes:
            best = np.inf
            for _ in range(repeats):
                o = hip.default_opts(tol=tol, rhs_tile=t, lanes=int(L), max_inner=probe_iters)
                st = hip.solve_probe(system, B, X, k, trans=trans, opts=o, seg=seg)
                if st["inner_iters"] > 0:
                    best = min(best, st["ms"])
            cost[f"{L}/{t}"] = best
    best = min(cost, key=cost.get)
    L, t = (int(v) for v in best.split("/"))
    choice = {"lanes": L, "rhs_tile": t}
    if use_cache and np.isfinite(cost[best]):
        store_choice(key, dict(choice, ms=cost[best]), path)
    return dict(choice, source="probe"), cost


def tune_rhs_tile(system, B, k, trans=False, seg=None, candidates=(16, 32, 64, 128, 256), probe_iters=64,
                  repeats=2):
    """RHS tile alone (automatic lanes): (best tile, 0 for one whole-width
    solve; {tile: best device ms of the capped probe})."""
    choice, cost = tune_solve(system, B, k, trans=trans, seg=seg, lanes=(0,), tiles=candidates,
                              probe_iters=probe_iters, repeats=repeats, use_cache=False)
    return choice["rhs_tile"], {int(c.split("/")[1]): v for c, v in cost.items()}
