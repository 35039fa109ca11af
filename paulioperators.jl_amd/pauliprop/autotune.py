"""Autotuner of the batched engine's tile shapes.

The engine has two tile parameters: the sweep tile (terms per flag/emit
tile: 512 or 2048) and the merge tile (outputs per merge workgroup: 512,
1024 or 2048, the last only for N <= 64).  For each key width W and live-size
class, the tuner times one layer of a real workload from the same starting
state (B copies of it in one batch) under every combination and keeps the
fastest.  The table (size class upper bounds with their tiles and the
measured times) is written as JSON; SparsePauliBatch.use_tuning(table)
then sets the tiles before each evolve call from the current largest live
size.
"""
from __future__ import annotations

import json
import os
import sys
import time
from typing import Callable, Dict, List, Optional

import numpy as np

from . import workloads as wl
from .batch import SparsePauliBatch

TUNING_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "tuning.json")

TILES = {1: [(512, 512), (512, 1024), (512, 2048), (2048, 1024), (2048, 2048)],
         2: [(512, 512), (512, 1024), (2048, 512), (2048, 1024)]}


def _states(W: int, log) -> List[tuple]:
    """(label, workload, layer index, arrays) of representative states: the
    36-qubit surrogate chain (W = 1) and the 121-qubit lattice and a 108-qubit
    grid instance of the batched set (W = 2), at several depths."""
    from .spv import SparsePauliVector
    out = []
    if W == 1:
        w = wl.surrogate_chain36_layers(wl.surrogate_thetas(1, seed=99)[0])
        stops = [2, 4, 6, 8]
    else:
        w = wl.lattice121(wl.theta_h_values()[3])
        stops = [3, 8, 14]
    v = SparsePauliVector(w.N, [((p.z, p.x), c) for p, c in w.observable])
    for li, L in enumerate(w.layers):
        if li in stops:
            out.append((f"{w.name}@{li}", w, li, v.arrays()))
            log(f"state {w.name} layer {li}: n = {len(v)}")
        if li >= stops[-1]:
            break
        v.evolve_(L.generators, L.angles, truncation=w.truncation)
        if w.channel:
            getattr(v, w.channel[0] + "_channel_")(w.channel[1])
    if W == 2:
        big = [x for x in wl.batched_instances(8, seed=7, max_depth=3) if x.N > 64 and x.meta["topology"] == "grid"][0]
        v = SparsePauliVector(big.N, [((p.z, p.x), c) for p, c in big.observable])
        for li in range(2):
            v.evolve_(big.layers[li].generators, big.layers[li].angles, truncation=big.truncation)
        out.append((f"{big.name}@2", big, 2, v.arrays()))
        log(f"state {big.name} layer 2: n = {len(v)}")
    return out


def _time_layer(w, li, arrays, copies: int, tile: int, mtile: int, reps: int) -> float:
    z, x, c = arrays
    best = float("inf")
    L = w.layers[li]
    for _ in range(reps):
        bt = SparsePauliBatch([w.N] * copies, [[]] * copies)
        for b in range(copies):
            bt.upload_arrays(b, z, x, c)
        bt.reserve(max(4 * len(c), 1 << 14))
        bt.set_tiles(tile, mtile)
        packed = bt.pack([L.generators] * copies, [L.angles] * copies)
        bt.sizes()
        t0 = time.perf_counter()
        bt.evolve_packed(packed, truncation=w.truncation)
        bt.sizes()
        best = min(best, time.perf_counter() - t0)
    return best


def autotune(path: str = TUNING_PATH, copies: int = 8, reps: int = 2,
             log: Optional[Callable[[str], None]] = None) -> Dict:
    log = log or (lambda m: None)
    table = {"copies": copies, "classes": {}}
    for W in (1, 2):
        rows = []
        for label, w, li, arrays in _states(W, log):
            n = len(arrays[2])
            times = {}
            for tile, mtile in TILES[W]:
                times[f"{tile}/{mtile}"] = _time_layer(w, li, arrays, copies, tile, mtile, reps)
            best = min(times, key=times.get)
            log(f"W={W} {label} n={n}: " + ", ".join(f"{k} {v * 1e3:.1f} ms" for k, v in times.items())
                + f" -> {best}")
            rows.append({"state": label, "n": n, "times_s": times, "best": best})
        rows.sort(key=lambda r: r["n"])
        # class upper bounds halfway (geometrically) between measured sizes
        classes = []
        for i, r in enumerate(rows):
            hi = int(np.sqrt(r["n"] * rows[i + 1]["n"])) if i + 1 < len(rows) else None
            t, m = (int(v) for v in r["best"].split("/"))
            classes.append({"n_max": hi, "tile": t, "merge_tile": m, "measured": r})
        table["classes"][str(W)] = classes
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as fh:
        json.dump(table, fh, indent=1)
    return table


def load(path: str = TUNING_PATH) -> Optional[Dict]:
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh)


def lookup(table: Dict, W: int, n: int):
    """(tile, merge_tile) for the largest live size n, or (0, 0) = engine default."""
    if not table:
        return 0, 0
    for c in table["classes"].get(str(W), []):
        if c["n_max"] is None or n <= c["n_max"]:
            return c["tile"], c["merge_tile"]
    return 0, 0


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="autotune the batched engine's tile shapes")
    ap.add_argument("--out", default=TUNING_PATH)
    ap.add_argument("--copies", type=int, default=8)
    ap.add_argument("--reps", type=int, default=2)
    a = ap.parse_args()
    autotune(a.out, a.copies, a.reps, log=lambda m: print(m, file=sys.stderr, flush=True))
