"""The batched workload: 256 independent circuits (seed 7; N ~ U{24..128},
depth ~ U{5..40}, chain or near-square grid, RZZ/RX/RZ angles U[0, π),
random single-site Z observable, |c| <= 1e-5) evolved together.

Instances are grouped by key width (N <= 64 and N > 64); each group is one
SparsePauliBatch, so every rotation step of the group is one fixed kernel
sequence for all its circuits, and each layer call is captured once as a
hipGraph and replayed while its shape holds.  Tile shapes come from the
autotuner's table when present.  Live-term budget: grid instances grow by
one to two orders of magnitude per layer under |c| <= 1e-5 (beyond 1e9 terms
within a few layers, more than one GPU holds), so an instance whose live
count passes `cap` terms is stopped at that point and reported; every other
instance runs to its full depth.  Steps are issued in chunks of `chunk`
rotations so that a growing instance is stopped close to the budget.
"""
from __future__ import annotations

import json
import math
import sys
import time
from typing import Callable, List, Optional

import numpy as np

from . import autotune
from . import workloads as wl
from ._lib import PPCounters
from .batch import SparsePauliBatch
from .spv import _pack


def _group_arrays(ws: List[wl.Workload], W: int):
    """Per layer li: packed steps (gz, gx, cv, sv, act) for the group, padded to
    the longest layer; gz/gx are shared by every layer of an instance."""
    B = len(ws)
    S = max(len(w.layers[0].generators) for w in ws)
    gz = np.zeros((S, B, W), dtype=np.uint64)
    gx = np.zeros((S, B, W), dtype=np.uint64)
    for b, w in enumerate(ws):
        for s, p in enumerate(w.layers[0].generators):
            gz[s, b] = _pack(p.z, W)
            gx[s, b] = _pack(p.x, W)
    depth = max(len(w.layers) for w in ws)
    layers = []
    for li in range(depth):
        cv = np.zeros((S, B))
        sv = np.zeros((S, B))
        act = np.zeros((S, B), dtype=np.uint8)
        for b, w in enumerate(ws):
            if li < len(w.layers):
                a = w.layers[li].angles
                k = len(a)
                cv[:k, b] = [math.cos(t) for t in a]
                sv[:k, b] = [math.sin(t) for t in a]
                act[:k, b] = 1
        layers.append((cv, sv, act))
    return gz, gx, layers


def run(count: int = 256, cap: int = 4_000_000, chunk: int = 64, graphs: bool = True, tuned: bool = True,
        max_depth: Optional[int] = None, log: Optional[Callable[[str], None]] = None):
    log = log or (lambda m: None)
    ws = wl.batched_instances(count, seed=7, max_depth=max_depth)
    table = autotune.load() if tuned else None
    groups = []
    for W, sel in ((1, [w for w in ws if w.N <= 64]), (2, [w for w in ws if w.N > 64])):
        if not sel:
            continue
        bt = SparsePauliBatch([w.N for w in sel], [[((p.z, p.x), c) for p, c in w.observable] for w in sel])
        bt.set_graphs(graphs)
        if table:
            bt.use_tuning(table)
        gz, gx, layers = _group_arrays(sel, bt.W)
        groups.append({"W": W, "ws": sel, "bt": bt, "gz": gz, "gx": gx, "layers": layers,
                       "alive": np.ones(len(sel), dtype=bool), "stopped_at": [None] * len(sel),
                       "max_n": np.zeros(len(sel), dtype=np.int64)})
    log(f"{count} instances: " + ", ".join(f"W={g['W']}: {len(g['ws'])}" for g in groups))
    swept = 0
    rotations = 0
    t0 = time.perf_counter()
    depth = max(len(w.layers) for w in ws)
    for li in range(depth):
        for g in groups:
            if li >= len(g["layers"]) or not g["alive"].any():
                continue
            cv, sv, act = g["layers"][li]
            act = act * g["alive"][None, :].astype(np.uint8)
            S = act.shape[0]
            for s0 in range(0, S, chunk):
                a = act[s0:s0 + chunk]
                if not a.any():
                    continue
                cnts = g["bt"].evolve_packed((g["gz"][s0:s0 + chunk], g["gx"][s0:s0 + chunk], cv[s0:s0 + chunk],
                                              sv[s0:s0 + chunk], np.ascontiguousarray(a)),
                                             truncation=g["ws"][0].truncation, counters=True)
                swept += sum(c.terms_swept for c in cnts)
                rotations += sum(c.rotations for c in cnts)
                n = np.array(g["bt"].sizes())
                g["max_n"] = np.maximum(g["max_n"], n)
                over = g["alive"] & (n > cap)
                for b in np.nonzero(over)[0]:
                    g["stopped_at"][b] = (li, s0 + chunk)
                g["alive"] &= ~over
                act[:, over] = 0
        if li % 5 == 0:
            log(f"layer {li}: {time.perf_counter() - t0:.1f} s, term-rotations {swept:.3e}, alive "
                + ", ".join(f"W={g['W']}: {int(g['alive'].sum())}" for g in groups))
    dt = time.perf_counter() - t0
    max_n = np.concatenate([g["max_n"] for g in groups])
    stopped = sum(sum(1 for s in g["stopped_at"] if s is not None) for g in groups)
    evs = []
    for g in groups:
        evs += g["bt"].expectation_values()
    return {
        "workload": f"batched: {count} instances seed 7, N U{{24..128}}, depth U{{5..40}}, chain/grid, RZZ/RX/RZ "
                    f"U[0,pi), single-site Z, |c|<=1e-5; live-term budget {cap} per instance",
        "seconds": dt, "term_rotations": swept, "rotations": rotations, "terms_per_s": swept / dt,
        "instances_completed": count - stopped, "instances_stopped_at_budget": stopped,
        "max_live_terms_p50_p90_max": [int(np.percentile(max_n, 50)), int(np.percentile(max_n, 90)),
                                       int(max_n.max())],
        "graphs": [g["bt"].graph_stats() for g in groups], "tuned": bool(table), "chunk": chunk,
    }


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=256)
    ap.add_argument("--cap", type=int, default=4_000_000)
    ap.add_argument("--chunk", type=int, default=64)
    ap.add_argument("--no-graphs", action="store_true")
    ap.add_argument("--no-tuning", action="store_true")
    ap.add_argument("--max-depth", type=int, default=None)
    a = ap.parse_args()
    r = run(a.count, a.cap, a.chunk, not a.no_graphs, not a.no_tuning, a.max_depth,
            log=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(r), flush=True)
