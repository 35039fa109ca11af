#!/usr/bin/env python3
"""Benchmark: Pauli terms/s per gate sweep on the 121-qubit lattice sweep.

Workload (one step): 121-qubit 11x11 open square lattice; 20 Trotter layers of
ZZ(θ_J = -π/2) on the 220 nearest-neighbour bonds then X(θ_h) on all sites;
θ_h swept over 16 values k (π/2)/15, k = 0..15, each an independent evolution;
dephasing p = 5e-4 per qubit after each layer; truncation |c| <= 1e-4 and
weight <= 12 after every rotation; observable (1/121) Σ_i Z_i; FP64.  The 16
initial operators are resident on the GPU when the timed region starts; the
16 evolutions run as one batch: each layer's 341 rotations in one resident
launch (the terms held in LDS by a cluster of workgroups per vector) where the
vectors fit, else one kernel sequence per rotation for all of them.  The
final term sets of the last timed step are checked against the committed
full-sweep fixture digests.

value = Σ over all rotations of the live terms entering the rotation / wall
time of the step.  Algorithmic bytes of a rotation = 40 B x (terms entering +
terms leaving it).  The JSON line carries that figure over the whole step
(roofline_rotation), per kernel over its HIP-event time (the resident layer
kernel and the per-step merge kernel; `roofline` is the one with more time),
the merge kernel where it is bandwidth-bound (256 circuits of a 36-qubit
chain with up to ~1e6 live terms each, 24 B per term), and two CPU
baselines: the single-threaded oracle (a C restatement of the reference
algorithm) on the same sweep, and the oracle in 16 processes on copies of the
16 sweeps; both final term sets are compared with the GPU's bit for bit.
HBM traffic per launch is not measured in this run (null); the rocprofv3
FETCH_SIZE/WRITE_SIZE figures are kept beside the kernel profiles.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "paulioperators.jl_amd"), os.path.join(ROOT, "oracle")]

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cpu() -> str:
    """The host CPU model (the CPU baselines run on one of its cores)."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def sweep_gpu(workloads, timing=True, graphs=False, digests=False, resident=1):
    """All θ_h instances in one batch: per layer one evolve! per instance, then
    the dephasing channel, then ⟨O⟩ of every instance.  Returns a dict:
    seconds (wall), counters (summed), evs, and when timing is on the HIP-event
    seconds and launch counts of the per-step merge kernel and of the resident
    layer kernel with the resident launches' share of the counters; digests:
    the final term sets' SHA-256 (taken after the clock stops)."""
    from pauliprop import PPCounters, SparsePauliBatch
    bt = SparsePauliBatch([w.N for w in workloads],
                          [[((p.z, p.x), c) for p, c in w.observable] for w in workloads])
    bt.reserve(1 << 18)
    bt.set_timing(timing)
    bt.set_graphs(graphs)
    bt.set_resident(resident, 0)
    tot = PPCounters()
    nl = len(workloads[0].layers)
    packed = [bt.pack([w.layers[li].generators for w in workloads], [w.layers[li].angles for w in workloads])
              for li in range(nl)]
    t0 = time.perf_counter()
    for li in range(nl):
        cnts = bt.evolve_packed(packed[li], truncation=workloads[0].truncation, counters=True)
        bt.channel_(workloads[0].channel[0], workloads[0].channel[1])
        for c in cnts:
            for f, _ in c._fields_:
                setattr(tot, f, getattr(tot, f) + getattr(c, f))
    evs = bt.expectation_values()
    dt = time.perf_counter() - t0
    out = {"seconds": dt, "counters": tot, "evs": evs, "resident_stats": bt.resident_stats()}
    if timing:
        out["merge"] = bt.merge_time()
        out["resident_seconds"] = bt.resident_time()
        out["resident_work"] = bt.resident_work()
    if digests:
        out["digests"] = [term_set_digest(*bt.arrays(b)) for b in range(len(workloads))]
    return out


def term_set_digest(z, x, c) -> str:
    """SHA-256 over the key words, then the coefficient bit patterns (the
    digest of the committed full-sweep fixture)."""
    import hashlib
    import numpy as np
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(z).tobytes())
    h.update(np.ascontiguousarray(x).tobytes())
    h.update(np.ascontiguousarray(c).view(np.uint64).tobytes())
    return h.hexdigest()


def merge_bandwidth_regime(count=256):
    """The merge kernel where it is bandwidth-bound: one batch of `count`
    circuits of the 36-qubit chain (10 layers, RZZ/RX angles from the seed-99
    stream, |c| <= 1e-6; up to ~1e6 live terms each).  Returns (algorithmic
    GB/s, average launch µs, bytes per launch, launches) from HIP events
    around every merge launch and the device merge counters (24 B per term)."""
    from pauliprop import SparsePauliBatch, workloads as wl
    from pauliprop.surrogate import _chain36_layer_arrays
    from pauliprop.truncation import CoeffTruncation
    th = [list(t) for t in wl.surrogate_thetas(count, seed=99)]
    obs = wl.surrogate_chain36_layers([0.0] * 20).observable
    bt = SparsePauliBatch([36] * count, [[((p.z, p.x), c) for p, c in obs]] * count)
    packed = [_chain36_layer_arrays(th, li, bt.W) for li in range(10)]
    bt.set_timing(True)
    mi = mo = 0
    for li in range(10):
        for c in bt.evolve_packed(packed[li], truncation=CoeffTruncation(1e-6), counters=True):
            mi += c.merge_in
            mo += c.merge_out
    t, n = bt.merge_time()
    b = 24 * (mi + mo)
    return b / t / 1e9, 1e6 * t / n, b / n, n


def complex_resident(theta_h=0.3, reps=2):
    """One 121-qubit lattice vector (θ_h = 0.3, all 20 layers with dephasing)
    as Float64 and as ComplexF64 (the reference's default element type), each
    with resident layer launches and on the per-step kernels: wall ms per run
    (best of `reps`), and whether the ComplexF64 run's real parts equal the
    Float64 run bit for bit (the observable's coefficients are real)."""
    import numpy as np
    from pauliprop import SparsePauliVector, workloads as wl
    w = wl.lattice121(theta_h)
    out = {"workload": f"lattice121 theta_h={theta_h}, 20 layers, one vector", "ms": {}}
    finals = {}
    for T in (float, complex):
        for mode in (1, 0):
            best = None
            for _ in range(reps):
                g = SparsePauliVector(w.N, [((p.z, p.x), T(c)) for p, c in w.observable], T=T).set_resident(mode, 0)
                t0 = time.perf_counter()
                for L in w.layers:
                    g.evolve_(L.generators, L.angles, truncation=w.truncation)
                    g.dephasing_channel_(w.channel[1])
                n = len(g)
                dt = time.perf_counter() - t0
                best = dt if best is None else min(best, dt)
            out["ms"][f"{T.__name__}_{'resident' if mode else 'per_step'}"] = best * 1e3
            finals[(T, mode)] = g.arrays()
    zf, xf, cf = finals[(float, 1)]
    zc, xc, cc = finals[(complex, 1)]
    out["terms"] = int(len(cf))
    out["complex_real_parts_equal_float64"] = bool(np.array_equal(zf, zc) and np.array_equal(xf, xc)
                                                   and np.array_equal(cf.view(np.uint64), cc.real.view(np.uint64))
                                                   and not np.any(cc.imag))
    return out


def algebra_commutator(reps=3, min_terms=20000, cpu=True):
    """[H, O] on the 121-qubit lattice (pp_algebra, complex coefficients):
    H = sum of Z_a Z_b over the 220 bonds + 0.5 sum of X_j (341 terms), O = the
    mean-Z observable evolved at theta_h = 0.3 on the GPU until it holds at
    least `min_terms` terms.  Returns the JSON record: pairs/s by wall time
    (host arrays in and out), the CPU baseline (the numpy oracle, stable
    order) and a bit-exact check of the GPU result against it."""
    import numpy as np
    from pauliprop import Pauli, SparsePauliVector, commutator
    from pauliprop import workloads as wl
    from pauliprop.spv import _pack
    N = 121
    w = wl.lattice121(0.3)
    v = SparsePauliVector(N, {(p.z, p.x): c for p, c in w.observable})
    layers = 0
    for L in w.layers:
        v.evolve_(L.generators, L.angles, truncation=w.truncation)
        layers += 1
        if len(v) >= min_terms:
            break
    oz, ox, oc = v.arrays()
    O = (oz, ox, oc.astype(np.complex128))
    hterms = [Pauli(N, Z=[a, b]) for a, b in wl.lattice_bonds(11, 11)] + [Pauli(N, X=[j]) for j in range(1, N + 1)]
    hz = np.array([_pack(p.z, 2) for p in hterms], dtype=np.uint64)
    hx = np.array([_pack(p.x, 2) for p in hterms], dtype=np.uint64)
    hc = np.array([1.0] * 220 + [0.5] * N, dtype=np.complex128)
    H = (hz, hx, hc)
    pairs = len(hc) * len(oc)
    out = commutator(N, H, O)  # warmup
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        out = commutator(N, H, O)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    rec = {"metric": "commutator pairs/s", "value": pairs / best, "unit": "pairs/s", "dtype": "c128",
           "seconds": best,
           "config": {"workload": f"[H, O] on the 121-qubit lattice: H = 220 ZZ bonds + 0.5 X on 121 sites, "
                                  f"O = mean Z evolved {layers} layers at theta_h = 0.3",
                      "pairs": pairs, "H_terms": len(hc), "O_terms": len(oc), "out_terms": len(out[2])},
           "cpu_baseline": None}
    if cpu:
        import algebra_oracle as ao
        t0 = time.perf_counter()
        ref = ao.product(H, O, ao.COMMUTATOR)
        cdt = time.perf_counter() - t0
        same = (np.array_equal(out[0], ref[0]) and np.array_equal(out[1], ref[1])
                and np.array_equal(out[2].view(np.uint64), ref[2].view(np.uint64)))
        assert same, "pp_algebra differs from the oracle"
        rec["cpu_baseline"] = {"value": pairs / cdt, "unit": "pairs/s", "cores": 1, "kind": "port", "host": host_cpu(),
                               "sample": f"the same commutator ({cdt:.1f} s, numpy oracle)"}
        rec["bitexact_vs_oracle"] = same
    return rec


def secondary_rows(cpu=True, layers=10):
    """The windowed evolution, device lookups and AdaptiveTruncation, each
    timed on the GPU and beside its CPU path, on the 64-qubit chain
    (RZZ/RX/RZ layers, depolarizing 1e-3, |c| <= 1e-6, weight <= 16) after
    `layers` layers (6.8e6 live terms):
      window    evolve! at window = 8 (and window = 1) over the first layers,
                against the oracle at window = 8 (the reference's quicksort
                order; <O> compared to 1e-10);
      lookup    getindex of 1e6 keys (half present) by device binary search,
                against numpy's searchsorted over the same sorted keys;
      adaptive  truncate!(v, AdaptiveTruncation(1e6)) by device radix select,
                against the oracle's sort-based restatement; terms compared
                to the bit."""
    import ctypes
    import numpy as np
    from pauliprop import AdaptiveTruncation, PPCounters, SparsePauliVector, compile_filter
    from pauliprop import _lib
    from pauliprop import workloads as wl
    w = wl.chain64()
    Ls = w.layers[:layers]
    obs = {(p.z, p.x): c for p, c in w.observable}
    recs = {}

    def gpu_evolve(window):
        v = SparsePauliVector(w.N, obs)
        swept = 0
        t0 = time.perf_counter()
        for L in Ls:
            cnt = PPCounters()
            v.evolve_(L.generators, L.angles, window=window, truncation=w.truncation, counters=cnt)
            v.depolarizing_channel_(w.channel[1])
            swept += cnt.terms_swept
        n = len(v)
        return v, swept, time.perf_counter() - t0, n

    gpu_evolve(8)  # warmup
    v8, sw8, t8, n8 = gpu_evolve(8)
    _, sw1, t1, n1 = gpu_evolve(1)
    ev8 = v8.expectation_value(0)
    rec = {"metric": "terms/s (windowed evolve!, window = 8)", "value": sw8 / t8, "unit": "terms/s",
           "seconds": t8, "window1_terms_per_s": sw1 / t1, "live_terms": n8,
           "config": {"workload": f"64-qubit chain, first {layers} layers, window 8", "term_rotations": sw8},
           "cpu_baseline": None}
    if cpu:
        import pyoracle as po
        o = po.OracleSPV(w.N, [(p.z, p.x, c) for p, c in w.observable])
        f = compile_filter(w.truncation)
        osw = 0
        t0 = time.perf_counter()
        for L in Ls:
            osw += o.evolve([(p.z, p.x) for p in L.generators], L.angles, window=8, strict=f).terms_swept
            o.channel(w.channel[0], w.channel[1])
        ct = time.perf_counter() - t0
        oev = o.expectation(0)
        assert abs(oev - ev8) <= 1e-10, ("windowed <O> vs oracle", ev8, oev)
        rec["cpu_baseline"] = {"value": osw / ct, "unit": "terms/s", "cores": 1, "kind": "port", "host": host_cpu(),
                               "sample": f"the same {layers} layers at window 8 ({ct:.1f} s)"}
        rec["expectation"] = ev8
        rec["expectation_oracle"] = oev
    recs["window"] = rec

    # lookups: 1e6 keys, half taken from the vector, half random
    z, x, c = v8.arrays()
    rng = np.random.default_rng(5)
    nq = 1_000_000
    pick = rng.integers(0, len(c), size=nq // 2)
    qz = np.concatenate([z[pick, 0], rng.integers(0, 2 ** 63, size=nq - nq // 2, dtype=np.uint64)])
    qx = np.concatenate([x[pick, 0], rng.integers(0, 2 ** 63, size=nq - nq // 2, dtype=np.uint64)])
    qz = np.ascontiguousarray(qz.reshape(-1, 1))
    qx = np.ascontiguousarray(qx.reshape(-1, 1))
    vals = np.zeros(nq)
    found = np.zeros(nq, dtype=np.int32)
    L = _lib.lib()
    args = (v8._h, _lib.u64p(qz), _lib.u64p(qx), nq, _lib.dp(vals), found.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    _lib.check(L.pp_lookup(*args))
    t0 = time.perf_counter()
    _lib.check(L.pp_lookup(*args))
    tl = time.perf_counter() - t0
    rec = {"metric": "lookups/s (getindex by device binary search)", "value": nq / tl, "unit": "keys/s",
           "seconds": tl, "config": {"workload": f"{nq} keys (half present) in {len(c)} live terms",
                                     "found": int(found.sum())}, "cpu_baseline": None}
    if cpu:
        keys = np.zeros(len(c), dtype=[("z", "<u8"), ("x", "<u8")])
        keys["z"], keys["x"] = z[:, 0], x[:, 0]
        q = np.zeros(nq, dtype=keys.dtype)
        q["z"], q["x"] = qz[:, 0], qx[:, 0]
        t0 = time.perf_counter()
        pos = np.searchsorted(keys, q)
        hit = pos < len(c)
        hit[hit] = keys[pos[hit]] == q[hit]
        cv = np.where(hit, c[np.minimum(pos, len(c) - 1)], 0.0)
        ct = time.perf_counter() - t0
        assert np.array_equal(hit, found.astype(bool)) and np.array_equal(cv[hit], vals[hit]), "lookup mismatch"
        rec["cpu_baseline"] = {"value": nq / ct, "unit": "keys/s", "cores": 1, "kind": "port", "host": host_cpu(),
                               "sample": f"numpy searchsorted, same keys ({ct:.2f} s)"}
    recs["lookup"] = rec

    # AdaptiveTruncation to 1e6 terms
    keep = 1_000_000
    va = SparsePauliVector(w.N, (z, x, c))
    len(va)
    t0 = time.perf_counter()
    va.truncate_(AdaptiveTruncation(keep, 1e-12))
    na = len(va)
    ta = time.perf_counter() - t0
    rec = {"metric": "terms/s (AdaptiveTruncation by radix select)", "value": len(c) / ta, "unit": "terms/s",
           "seconds": ta, "config": {"workload": f"{len(c)} live terms truncated to {keep}", "kept": na},
           "cpu_baseline": None}
    if cpu:
        import pyoracle as po
        o2 = po.OracleSPV(w.N)
        zc, xc, cc = (np.ascontiguousarray(a) for a in (z, x, c))
        po.lib().or_load(o2.h, po._u64p(zc), po._u64p(xc), po._dp(cc), len(cc))  # grows, sorts (already sorted)
        t0 = time.perf_counter()
        o2.adaptive(keep, 1e-12)
        ct = time.perf_counter() - t0
        gz, gx, gc = va.arrays()
        oz, ox, oc = o2.arrays()
        same = np.array_equal(gz, oz) and np.array_equal(gx, ox) and np.array_equal(gc.view(np.uint64), oc.view(np.uint64))
        assert same, "adaptive truncation differs from the oracle"
        rec["cpu_baseline"] = {"value": len(c) / ct, "unit": "terms/s", "cores": 1, "kind": "port", "host": host_cpu(),
                               "sample": f"oracle or_adaptive (sort-based), same vector ({ct:.2f} s)"}
        rec["bitexact_vs_oracle"] = same
    recs["adaptive"] = rec
    return recs


def _oracle_instances(args):
    """One worker of the all-cores CPU baseline: the oracle on the batched
    instances with index = worker (mod workers), same layers, chunks and stop
    rule, until the deadline.  Returns (term-rotations, seconds, instances)."""
    worker, workers, max_depth, cap, budget_s = args
    import pyoracle as po
    from pauliprop import compile_filter
    from pauliprop import workloads as wl
    ws = wl.batched_instances(256, seed=7, max_depth=max_depth)
    swept = done = 0
    t0 = time.perf_counter()
    for w in ws[worker::workers]:
        o = po.OracleSPV(w.N, [(p.z, p.x, c) for p, c in w.observable])
        f = compile_filter(w.truncation)
        stop = False
        for L in w.layers:
            g = [(p.z, p.x) for p in L.generators]
            for s0 in range(0, len(g), 64):
                swept += o.evolve(g[s0:s0 + 64], L.angles[s0:s0 + 64], strict=f).terms_swept
                if len(o) > cap:
                    stop = True
                    break
            if stop:
                break
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    return swept, time.perf_counter() - t0, done


def batched_with_cpu(max_depth=6, cap=1_000_000, budget_s=20.0, cpu=True, workers=16):
    """The batched workload (256 instances, seed 7) over its first `max_depth`
    layers on the GPU (pauliprop.batched_bench.run: two key-width batches,
    captured graphs, autotuned tiles, instances stopped past `cap` live
    terms), and the oracle on a bounded sample of the same instances: one
    instance at a time in seed order, the same 64-rotation chunks and the same
    stop rule, until `budget_s` seconds of CPU work."""
    from pauliprop import batched_bench, compile_filter
    from pauliprop import workloads as wl
    rec = batched_bench.run(256, cap, 64, True, True, max_depth, log=log)
    rec = {"metric": "terms/s (batched, first layers)", "value": rec["terms_per_s"], "unit": "terms/s", **rec,
           "cpu_baseline": None}
    if cpu:
        import pyoracle as po
        ws = wl.batched_instances(256, seed=7, max_depth=max_depth)
        swept = 0
        done = 0
        t0 = time.perf_counter()
        for w in ws:
            o = po.OracleSPV(w.N, [(p.z, p.x, c) for p, c in w.observable])
            f = compile_filter(w.truncation)
            stop = False
            for L in w.layers:
                g = [(p.z, p.x) for p in L.generators]
                for s0 in range(0, len(g), 64):
                    swept += o.evolve(g[s0:s0 + 64], L.angles[s0:s0 + 64], strict=f).terms_swept
                    if len(o) > cap:
                        stop = True
                        break
                if stop:
                    break
            done += 1
            if time.perf_counter() - t0 > budget_s:
                break
        ct = time.perf_counter() - t0
        rec["cpu_baseline"] = {"value": swept / ct, "unit": "terms/s", "cores": 1, "kind": "port", "host": host_cpu(),
                               "sample": f"the first {done} of 256 instances, same layers and stop rule "
                                         f"({swept} term-rotations, {ct:.1f} s)"}
        # all cores: independent instances per worker process
        from concurrent.futures import ProcessPoolExecutor
        import multiprocessing as mp
        with ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
            res = list(ex.map(_oracle_instances, [(k, workers, max_depth, cap, budget_s) for k in range(workers)]))
        tot = sum(r[0] for r in res)
        span = max(r[1] for r in res)
        rec["cpu_baseline_all_cores"] = {"value": tot / span, "unit": "terms/s", "cores": workers, "kind": "port",
                                         "host": host_cpu(),
                                         "sample": f"{sum(r[2] for r in res)} instances over {workers} worker "
                                                   f"processes, one instance at a time each ({tot} term-rotations, "
                                                   f"{span:.1f} s)"}
    return rec


def sweep_cpu(workloads, budget_s):
    """The oracle on the sweep, one θ_h at a time in order, until budget_s
    seconds.  Returns (terms/s, θ_h done, term-rotations, seconds, <O> list,
    final digests); the digests are taken outside the timed loop."""
    import pyoracle as po
    from pauliprop import compile_filter
    swept = 0
    dt = 0.0
    evs, digs = [], []
    for w in workloads:
        t0 = time.perf_counter()
        o = po.OracleSPV(w.N, [(p.z, p.x, c) for p, c in w.observable])
        f = compile_filter(w.truncation)
        for L in w.layers:
            c = o.evolve([(p.z, p.x) for p in L.generators], L.angles, strict=f)
            o.channel(w.channel[0], w.channel[1])
            swept += c.terms_swept
        evs.append(o.expectation(0))
        dt += time.perf_counter() - t0
        digs.append(term_set_digest(*o.arrays()))
        if dt > budget_s:
            break
    return swept / dt, len(evs), swept, dt, evs, digs


_SWEEP_WS = None


def _sweep_worker_init(barrier):
    """All-cores CPU baseline worker: build the 16 workloads and load the
    oracle, then wait for every worker so the timed span starts together."""
    global _SWEEP_WS
    import pyoracle  # noqa: F401
    from pauliprop import workloads as wl
    _SWEEP_WS = [wl.lattice121(t) for t in wl.theta_h_values()]
    barrier.wait(timeout=300)


def _sweep_job(k):
    """One full 20-layer sweep of θ_h index k on the oracle.  Returns
    (k, term-rotations, wall start, wall end, final digest)."""
    import pyoracle as po
    from pauliprop import compile_filter
    w = _SWEEP_WS[k]
    t0 = time.time()
    o = po.OracleSPV(w.N, [(p.z, p.x, c) for p, c in w.observable])
    f = compile_filter(w.truncation)
    swept = 0
    for L in w.layers:
        swept += o.evolve([(p.z, p.x) for p in L.generators], L.angles, strict=f).terms_swept
        o.channel(w.channel[0], w.channel[1])
    t1 = time.time()
    return k, swept, t0, t1, term_set_digest(*o.arrays())


def sweep_cpu_all_cores(reps, workers=16):
    """The sweep on every host core: `reps` copies of the 16 θ_h sweeps as
    independent jobs handed to `workers` single-threaded oracle processes
    (one job at a time each).  Returns (terms/s over the span from the first
    job's start to the last job's end, jobs, term-rotations, span, digests by
    θ_h index)."""
    from concurrent.futures import ProcessPoolExecutor
    import multiprocessing as mp
    ctx = mp.get_context("spawn")
    barrier = ctx.Barrier(workers)
    jobs = [k for _ in range(reps) for k in range(16)]
    with ProcessPoolExecutor(max_workers=workers, mp_context=ctx, initializer=_sweep_worker_init,
                             initargs=(barrier,)) as ex:
        res = list(ex.map(_sweep_job, jobs, chunksize=1))
    tot = sum(r[1] for r in res)
    span = max(r[3] for r in res) - min(r[2] for r in res)
    digs = {}
    for k, _, _, _, d in res:
        assert digs.setdefault(k, d) == d, "oracle runs of one θ_h differ"
    return tot / span, len(jobs), tot, span, digs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-budget", type=float, default=30.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-workers", type=int, default=16, help="processes of the all-cores CPU baseline")
    ap.add_argument("--cpu-all-seconds", type=float, default=6.0,
                    help="target wall seconds of the all-cores CPU baseline")
    ap.add_argument("--graphs", type=int, default=0, help="replay hipGraph-captured layer sequences (1/0)")
    ap.add_argument("--resident", type=int, default=1,
                    help="resident layer launches: 1 where the vectors fit (default), 0 the per-step kernels only")
    ap.add_argument("--batched", action="store_true",
                    help="run only the batched workload's first 6 layers with its CPU baseline")
    ap.add_argument("--rows", action="store_true",
                    help="run only the windowed-evolution, lookup and AdaptiveTruncation measurements")
    ap.add_argument("--algebra", action="store_true",
                    help="run only the Pauli-sum product measurement ([H, O] on the 121-qubit lattice)")
    ap.add_argument("--complex", action="store_true",
                    help="run only the Float64 / ComplexF64 single-vector comparison (resident and per-step)")
    ap.add_argument("--bandwidth", action="store_true",
                    help="run only the bandwidth-regime merge measurement (256 circuits of the 36-qubit chain)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a 1 GiB device-to-device copy (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.empty(1 << 30, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": 1 << 30, "launches": 3}))
        return
    if args.batched:
        print(json.dumps(batched_with_cpu(cpu=not args.no_cpu)), flush=True)
        return
    if args.complex:
        print(json.dumps(complex_resident()), flush=True)
        return
    if args.bandwidth:
        gbs, us, nbytes, launches = merge_bandwidth_regime()
        print(json.dumps({"workload": "256 circuits of the 36-qubit chain, 10 layers, |c| <= 1e-6",
                          "merge_gbs": gbs, "merge_avg_launch_us": us, "merge_bytes_per_launch": nbytes,
                          "merge_launches": launches, "frac_of_8tbs": gbs / HBM_PEAK_GBS}), flush=True)
        return
    if args.rows:
        print(json.dumps(secondary_rows(cpu=not args.no_cpu)), flush=True)
        return
    if args.algebra:
        print(json.dumps(algebra_commutator(cpu=not args.no_cpu)), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: measures one GPU only; --gpus must be 1 (the sweep's 16 theta_h vectors "
              "would shard across GPUs as independent vectors, with no collective on the step path)",
              file=sys.stderr)
        sys.exit(2)

    from pauliprop import PPCounters
    from pauliprop import workloads as wl

    ths = wl.theta_h_values()
    workloads = [wl.lattice121(t) for t in ths]
    S = 40  # bytes per term: 2 x 16 B keys + 8 B coefficient (65 <= N <= 128)
    with open(os.path.join(ROOT, "tests", "golden", "lattice121_full.json")) as fh:
        fixture = json.load(fh)

    graphs = bool(args.graphs)
    for i in range(args.warmup):
        log(f"warmup {i}")
        sweep_gpu(workloads, timing=False, graphs=graphs, resident=args.resident)

    tot_time = 0.0
    agg = PPCounters()
    evs = digs = None
    for i in range(args.steps):
        log(f"step {i}")
        r = sweep_gpu(workloads, timing=False, graphs=graphs, digests=(i == args.steps - 1), resident=args.resident)
        cnt = r["counters"]
        log(f"  {r['seconds']:.3f} s, {cnt.terms_swept} term-rotations, {cnt.rotations} rotations, "
            f"resident launches/fallbacks/aborts {r['resident_stats']}")
        tot_time += r["seconds"]
        for f, _ in cnt._fields_:
            setattr(agg, f, getattr(agg, f) + getattr(cnt, f))
        evs, digs = r["evs"], r.get("digests", digs)
    # every final term set against the committed full-sweep fixture
    bad = [k for k, (d, fx) in enumerate(zip(digs, fixture)) if d != fx["digest"]]
    assert not bad, f"final term sets differ from the fixture at theta_h indices {bad}"
    log("  final term sets of all 16 theta_h match the fixture digests")

    # kernel rooflines: one more (untimed) sweep with HIP events around every
    # resident layer launch and every per-step merge launch
    pr = sweep_gpu(workloads, timing=True, graphs=False, resident=args.resident)
    assert pr["evs"] == evs, "profiled sweep differs from the timed one"
    pc = pr["counters"]
    t_merge, n_merge = pr["merge"]
    t_res = pr["resident_seconds"]
    n_res = pr["resident_stats"][0]
    r_rot, r_swept, r_out, r_rebal = pr["resident_work"]
    log(f"  profiled sweep: resident kernel {t_res:.3f} s over {n_res} launches ({r_rot} rotations, "
        f"{r_rebal} chunk rebalances), "
        f"per-step merge kernel {t_merge:.3f} s over {n_merge} launches")

    def kernel_roofline(kernel, nbytes, seconds, launches, rotations, note):
        achieved = nbytes / seconds / 1e9 if seconds > 0 else None
        return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": None, "kernel": kernel,
                "launches": launches, "rotations": rotations, "seconds": seconds,
                "avg_launch_us": 1e6 * seconds / max(launches, 1), "bytes_per_launch": nbytes / max(launches, 1),
                "bytes": note}

    note = "S x (terms entering + terms leaving each rotation), S = 40 B"
    k_res = kernel_roofline("kr_layer<2>", S * (r_swept + r_out), t_res, n_res, r_rot, note)
    k_merge = kernel_roofline("kb_merge<2,512>", S * (pc.terms_swept - r_swept + pc.merge_out - r_out), t_merge,
                              n_merge, pc.rotations - r_rot, note)
    dominant, other = (k_res, k_merge) if t_res >= t_merge else (k_merge, k_res)
    if not other["launches"]:
        other = None  # every rotation ran in the other kernel
    step_bytes = S * (agg.terms_swept + agg.merge_out) / args.steps
    step_s = tot_time / args.steps
    rotation = {"bound": "hbm", "achieved": step_bytes / step_s / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": step_bytes / step_s / 1e9 / HBM_PEAK_GBS, "bytes_per_step": step_bytes,
                "bytes": note + ", over the whole step's wall time"}

    log("merge kernel in the bandwidth-bound regime (36-qubit chain, 256 circuits)")
    bw_gbs, bw_us, bw_bytes, bw_n = merge_bandwidth_regime()
    log(f"  {bw_gbs:.0f} GB/s algorithmic, {bw_us:.1f} us per launch over {bw_n} launches")

    cpu = cpu_all = None
    if not args.no_cpu:
        log("cpu baseline (oracle, 1 thread)")
        cv, done, cs, cdt, cevs, cdigs = sweep_cpu(workloads, args.cpu_budget)
        for k, (a, b) in enumerate(zip(evs, cevs)):
            assert abs(a - b) <= 1e-12, ("<O> mismatch vs oracle", k, a, b)
        for k, d in enumerate(cdigs):
            assert d == digs[k], f"final term set differs from the oracle at theta_h index {k}"
        cpu = {"value": cv, "unit": "terms/s", "cores": 1, "kind": "port", "host": host_cpu(),
               "sample": f"the same sweep, first {done} of 16 theta_h values ({cs} term-rotations, {cdt:.1f} s); "
                         f"final term sets bit-identical to the GPU's"}
        per_sweep = cdt * 16 / done
        reps = max(1, int(round(args.cpu_all_seconds * args.cpu_workers / per_sweep)))
        log(f"cpu baseline (oracle, {args.cpu_workers} processes, {reps} x 16 sweeps)")
        av, jobs, at, span, adigs = sweep_cpu_all_cores(reps, args.cpu_workers)
        for k, d in adigs.items():
            assert d == digs[k], f"all-cores oracle differs from the GPU at theta_h index {k}"
        cpu_all = {"value": av, "unit": "terms/s", "cores": args.cpu_workers, "kind": "port", "host": host_cpu(),
                   "sample": f"{jobs} full theta_h sweeps ({reps} x the 16 values) over {args.cpu_workers} "
                             f"single-threaded oracle processes, one sweep at a time each ({at} term-rotations, "
                             f"{span:.1f} s from the first start to the last end)"}

    value = agg.terms_swept / tot_time
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
        "config": {"workload": "lattice121_sweep: 121-qubit 11x11 open lattice, 20 layers ZZ(-pi/2) on 220 bonds "
                               "+ X(theta_h) on 121 sites, 16 theta_h in [0, pi/2], dephasing 5e-4, |c|<=1e-4, "
                               "weight<=12, observable mean Z",
                   "rotations": agg.rotations // args.steps, "term_rotations": agg.terms_swept // args.steps,
                   "graphs": graphs, "resident": args.resident, "expectation_by_theta_h": evs,
                   "final_terms_by_theta_h": [f["n_by_layer"][-1] for f in fixture],
                   "final_term_sets_match_fixture": True},
        "roofline": dominant,
        "roofline_other_kernel": other,
        "roofline_rotation": rotation,
        "roofline_bandwidth_regime": {"bound": "hbm", "achieved": bw_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                      "frac": bw_gbs / HBM_PEAK_GBS, "kernel": "kb_merge<1,2048>",
                                      "workload": "36-qubit chain, 10 layers, 256 circuits in one batch, |c|<=1e-6",
                                      "launches": bw_n, "avg_launch_us": bw_us, "bytes_per_launch": bw_bytes},
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
