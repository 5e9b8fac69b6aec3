"""Autotuner of the P2P kernel's launch and work-item shape.

Times candidate tunings of one prepared list on the device and keeps the
fastest.  The knobs are the ones pp2p_tuning exposes:
  ipl_max      i-particles per lane an item may use (1..3); with it the item
               builder picks G phases x IPL slots per item
  split_pairs  size of a queue item (pair-interactions) = the j-range a wave
               owns before it fetches the next item
  waves_per_cu persistent waves per CU of the queue form (0 = occupancy max)
  law          the cubic LOGTAB table, the linear table (LINEAR) or the
               table-free POLY law; a law other than the cubic is kept only
               when its forces stay within LAW_GAP_P99 of the cubic's
Tuned values are kept in photons_p2p/tuned.json, keyed by a signature of
the workload shape (particle, leaf, task and pair counts in half-octave
buckets); `lookup` returns the entry for a shape, `tune` measures a new one.
"""
from __future__ import annotations

import json
import os
import time

from . import Context, EXEC_QUEUE, ForceLaw, LAW_LINEAR, LAW_POLY, LAW_TABLE, P2PError

TUNED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned.json")

GRID = {
    "ipl_max": [1, 2, 3],
    "split_pairs": [0, 1 << 15, 1 << 16, 1 << 17, 1 << 18, 1 << 19, 1 << 20, 1 << 22],
    "waves_per_cu": [0, 12, 8],
    "law": [LAW_TABLE, LAW_LINEAR, LAW_POLY],
}


def signature(ntask: int, npart: int, nleaf: int, npairs: int) -> str:
    b = lambda x: int(round(2 * __import__("math").log2(max(x, 1))))  # noqa: E731
    return f"p{b(npart)}-l{b(nleaf)}-t{b(ntask)}-n{b(npairs)}"


# A law other than the cubic table is kept only if its accelerations differ
# from the cubic's by at most this much (p99 over the list's target
# particles, relative): half the 1e-5 force-error budget, so that the cubic's
# own error (1-3e-6) plus the difference stays inside it.
LAW_GAP_P99 = 5e-6


def law_gap(ctx: Context, law: ForceLaw, cfg: dict) -> float:
    """p99 of |a_cfg - a_cubic| / |a_cubic| over the particles that receive a
    force, both evaluated on the device for the context's list."""
    import numpy as np
    apply(ctx, law, {**cfg, "law": LAW_TABLE})
    ctx.zero_acc()
    ctx.compute(EXEC_QUEUE)
    ref = ctx.acc()
    apply(ctx, law, cfg)
    ctx.zero_acc()
    ctx.compute(EXEC_QUEUE)
    a = ctx.acc()
    mag = np.linalg.norm(ref, axis=1)
    sel = mag > 0
    if not np.any(sel):
        return 0.0
    return float(np.percentile(np.linalg.norm(a[sel] - ref[sel], axis=1) / mag[sel], 99))


def _time(ctx: Context, steps: int) -> float:
    """ms per step: at least `steps` steps and at least about 50 ms of device
    time, so that short steps (a fraction of a millisecond) are not timed
    from a handful of launches."""
    ctx.time_steps(EXEC_QUEUE, 1)                       # warm (builds the queue for this tuning)
    est, _ = ctx.time_steps(EXEC_QUEUE, 1)
    n = max(steps, min(400, int(50.0 / max(est, 1e-3))))
    tot, _ = ctx.time_steps(EXEC_QUEUE, n)
    return tot / n


def tune(ctx: Context, law: ForceLaw, steps: int = 3, grid: dict | None = None, log=None,
         budget_s: float = 120.0) -> dict:
    """Coordinate search over `grid` (one knob at a time, in the order above,
    starting from the defaults).  The context must hold particles and a list.
    Returns {"best": tuning, "ms": ms_per_step, "trials": [...]}; the context
    is left with the best tuning applied."""
    g = dict(GRID if grid is None else grid)
    best = {"ipl_max": 0, "split_pairs": 0, "waves_per_cu": 0, "law": law.law}
    trials = []
    t_start = time.time()

    def set_cfg(cfg):
        apply(ctx, law, cfg)

    set_cfg(best)
    best_ms = _time(ctx, steps)
    trials.append({**best, "ms": best_ms})
    for knob in ("ipl_max", "split_pairs", "waves_per_cu", "law"):
        for v in g.get(knob, []):
            if v == best[knob] or time.time() - t_start > budget_s:
                continue
            cfg = {**best, knob: v}
            try:
                set_cfg(cfg)
            except P2PError:                        # e.g. a law whose range excludes this softening
                set_cfg(best)
                continue
            ms = _time(ctx, steps)
            trial = {**cfg, "ms": ms}
            if ms < best_ms and knob == "law" and v != LAW_TABLE:
                gap = law_gap(ctx, law, cfg)
                trial["law_gap_p99"] = gap
                if gap > LAW_GAP_P99:
                    trial["rejected"] = "accuracy"
                    ms = float("inf")
                set_cfg(cfg)
            trials.append(trial)
            if log:
                log(f"[autotune] {trial}")
            if ms < best_ms:
                best, best_ms = cfg, ms
    set_cfg(best)
    return {"best": best, "ms": best_ms, "trials": trials}


def load_table(path: str = TUNED) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save_table(entries: dict, path: str):
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(entries, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def lookup(sig: str, path: str = TUNED) -> dict | None:
    """Tuned knobs for a workload shape, or None."""
    e = load_table(path).get(sig)
    return None if e is None else e.get("best")


def apply(ctx: Context, law: ForceLaw, cfg: dict):
    ctx.set_law(ForceLaw(eps=law.eps, rs=law.rs, mass=law.mass, ucut=law.ucut, softening=law.softening,
                         law=cfg.get("law", law.law), frame=law.frame, box=law.box))
    ctx.set_tuning(waves_per_cu=cfg.get("waves_per_cu", 0), split_pairs=cfg.get("split_pairs", 0),
                   ipl_max=cfg.get("ipl_max", 0))


__all__ = ["tune", "lookup", "apply", "signature", "GRID", "load_table", "save_table"]
