#!/usr/bin/env python3
"""Benchmark: Pauli terms/s per gate sweep on the 64-qubit chain workload.

Workload (one step): 64-qubit open chain, 30 layers; each layer RZZ on the 63
bonds, then RX and RZ on the 64 sites, angles i.i.d. U[0, π) from seed 1234;
depolarizing p = 1e-3 after each layer; truncation |c| <= 1e-6 and weight
<= 16 after every rotation; observable Z_32; FP64 coefficients.  The state
starts resident on the GPU; the timed region is the full evolution.

value = sum over rotations of the live terms entering the rotation / wall
time of the step.  The JSON line also carries the merge kernel's roofline
(algorithmic bytes = 24 B x (terms in + children + terms out) per launch,
divided by its HIP-event time) and a CPU baseline: the single-threaded
oracle on a bounded sample of the same workload.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "paulioperators.jl_amd"), os.path.join(ROOT, "oracle")]

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_workload(w, layers, counters_cls, spv_cls):
    obs = [((p.z, p.x), c) for p, c in w.observable]
    v = spv_cls(w.N, obs)
    tot = counters_cls()
    t0 = time.perf_counter()
    for li in range(layers):
        L = w.layers[li]
        cnt = counters_cls()
        v.evolve_(L.generators, L.angles, truncation=w.truncation, counters=cnt)
        if w.channel:
            kind, p = w.channel
            getattr(v, kind + "_channel_")(p)
        for f, _ in cnt._fields_:
            setattr(tot, f, getattr(tot, f) + getattr(cnt, f))
        log(f"  layer {li}: n={len(v)} swept={cnt.terms_swept} t={time.perf_counter() - t0:.2f}s")
    ev = v.expectation_value(0)
    return time.perf_counter() - t0, tot, ev, len(v)


def cpu_baseline(w, budget_s=20.0):
    """Single-threaded oracle on the first layers of the same workload, stopping
    once the budget is spent (whole layers only)."""
    import pyoracle as po
    from pauliprop import compile_filter
    o = po.OracleSPV(w.N, [(p.z, p.x, c) for p, c in w.observable])
    f = compile_filter(w.truncation)
    swept = 0
    t0 = time.perf_counter()
    layers = 0
    for L in w.layers:
        c = o.evolve([(p.z, p.x) for p in L.generators], L.angles, strict=f)
        if w.channel:
            kind, p = w.channel
            o.channel(kind, p)
        swept += c.terms_swept
        layers += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return swept / dt, layers, swept, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--layers", type=int, default=30)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU; runs on more than one GPU are not part of it "
              "(the sharding design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from pauliprop import PPCounters, SparsePauliVector
    from pauliprop import workloads as wl

    w = wl.chain64(n_layers=30)
    layers = min(args.layers, 30)
    S = 24  # bytes per term: 2 x 8 B keys + 8 B coefficient (N <= 64)

    for i in range(args.warmup):
        log(f"warmup {i}")
        run_workload(w, min(layers, 8), PPCounters, SparsePauliVector)

    tot_time = 0.0
    agg = PPCounters()
    ev = None
    nfin = 0
    for i in range(args.steps):
        log(f"step {i}")
        dt, cnt, ev, nfin = run_workload(w, layers, PPCounters, SparsePauliVector)
        tot_time += dt
        for f, _ in cnt._fields_:
            setattr(agg, f, getattr(agg, f) + getattr(cnt, f))

    value = agg.terms_swept / tot_time
    merge_bytes = S * (agg.merge_in + agg.merge_out)
    achieved = merge_bytes / agg.t_merge / 1e9 if agg.t_merge > 0 else None
    rot_bytes = S * (agg.terms_swept + agg.merge_out)
    rot_gbs = rot_bytes / (agg.t_rotate + agg.t_merge) / 1e9 if agg.t_merge > 0 else None

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle, 1 thread)")
        cv, cl, cs, cdt = cpu_baseline(w, args.cpu_budget)
        cpu = {"value": cv, "unit": "terms/s", "cores": 1, "kind": "port",
               "sample": f"chain64 layers 1-{cl} from the start ({cs} term-rotations, {cdt:.1f} s)"}

    out = {
        "metric": "Pauli terms/s per gate sweep",
        "value": value,
        "unit": "terms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * tot_time / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "chain64_depol: 64-qubit open chain, 30 layers RZZ+RX+RZ, seed 1234, "
                               "depolarizing 1e-3, |c|<=1e-6, weight<=16, observable Z_32",
                   "layers": layers, "rotations": agg.rotations, "final_terms": nfin, "expectation": ev},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": None,
                     "kernel": "k_merge<1>", "rotation_gbs": rot_gbs},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
