"""⟨O⟩ evaluations per second on the 36-qubit surrogate workload: plain
one-at-a-time GPU propagation, batched GPU propagation, and the trained
ensemble with its disagreement gate (batched propagation for the queries it
rejects), with |Δ⟨O⟩| judged against exact propagation on held-out θ.
Prints one JSON object."""
from __future__ import annotations

import json
import sys
import time

import numpy as np

from . import surrogate as sg


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run(n_plain: int = 8, n_heldout: int = 4096, epochs: int = 60, n_query: int = 1_000_000, tol: float = 1e-3,
        encoding: str = "trig"):
    lab = np.load(sg.LABELS_PATH)["labels"]
    tr, ho = sg.training_thetas()
    ytr, yho = lab[:len(tr)], lab[len(tr):]
    ho, yho = ho[:n_heldout], yho[:n_heldout]
    res = {"workload": "chain36 surrogate: 36-qubit chain, 10 layers (RZZ, RX shared per layer), theta in "
                       "[-pi,pi]^20 seed 99, observable Z_18, |c|<=1e-6", "tol": tol}
    # plain: one circuit at a time
    sg.chain36_plain_expectation(ho[0])
    t0 = time.perf_counter()
    plain = [sg.chain36_plain_expectation(t) for t in ho[:n_plain]]
    t_plain = time.perf_counter() - t0
    res["plain_evals_per_s"] = n_plain / t_plain
    res["plain_max_abs_err"] = float(np.max(np.abs(np.array(plain) - yho[:n_plain])))
    log(f"plain: {n_plain} in {t_plain:.2f} s")
    # batched exact, on the held-out θ the gate evaluates below
    t0 = time.perf_counter()
    bvals = sg.chain36_batch_expectations(ho, batch=256)
    t_b = time.perf_counter() - t0
    res["batched_evals_per_s"] = len(ho) / t_b
    res["batched_max_abs_err"] = float(np.max(np.abs(bvals - yho)))
    log(f"batched: {len(ho)} in {t_b:.2f} s")
    # training (the last 4,096 training θ are kept back to calibrate the gate)
    ncal = 4096
    res["encoding"] = encoding
    ens = sg.MLPEnsemble(E=5, din=sg.feature_dim(encoding), hidden=512, layers=3, seed=0)
    hist, t_train = ens.fit(sg.features(tr[:-ncal], encoding), ytr[:-ncal], epochs=epochs, batch=1024, lr=1e-3,
                            lr_min=1e-5, log=log)
    res["train_seconds"] = t_train
    res["train_epochs"] = epochs
    res["train_final_loss"] = [float(v) for v in hist[-1]]
    steps = epochs * ((len(tr) - ncal) // 1024)
    flop_step = 5 * 1024 * 2 * (64 * 512 + 2 * 512 * 512) * 3
    res["train_tflops"] = steps * flop_step / t_train / 1e12
    # ensemble quality on held-out
    mean, sd, _, _ = ens.predict(sg.features(ho, encoding))
    err = np.abs(mean - yho)
    res["ensemble_rmse_heldout"] = float(np.sqrt(np.mean(err ** 2)))
    res["ensemble_std_median"] = float(np.median(sd))
    # gated evaluation of the held-out θ
    gate = sg.GatedSurrogate(ens, exact=lambda th: sg.chain36_batch_expectations(th, batch=256), tol=tol,
                             encoding=encoding)
    res["gate_calibration"] = gate.calibrate(tr[-ncal:], ytr[-ncal:])
    t0 = time.perf_counter()
    vals, fall, tm = gate.evaluate(ho)
    t_g = time.perf_counter() - t0
    gerr = np.abs(vals - yho)
    res["gated_evals_per_s"] = len(ho) / t_g
    res["gated_served_by_network"] = int((~fall).sum())
    res["gated_max_abs_err"] = float(gerr.max())
    res["gated_max_abs_err_served"] = float(gerr[~fall].max()) if (~fall).any() else None
    res["gated_timing"] = tm
    log(f"gated: {len(ho)} in {t_g:.2f} s, served {(~fall).sum()}")
    # inference throughput on the query set
    q = sg.query_thetas(n_query)
    t0 = time.perf_counter()
    _, qsd, _, dev = ens.predict(sg.features(q, encoding))
    t_q = time.perf_counter() - t0
    res["inference_queries"] = n_query
    res["inference_evals_per_s_wall"] = n_query / t_q
    res["inference_evals_per_s_device"] = n_query / dev
    res["inference_tflops_device"] = n_query * 5 * 2 * (64 * 512 + 2 * 512 * 512 + 512) / dev / 1e12
    res["query_fraction_served"] = float(np.mean(gate.serves(qsd)))
    res["speedup_gated_vs_plain"] = res["gated_evals_per_s"] / res["plain_evals_per_s"]
    res["speedup_batched_vs_plain"] = res["batched_evals_per_s"] / res["plain_evals_per_s"]
    res["speedup_gated_vs_batched_same_theta"] = res["gated_evals_per_s"] / res["batched_evals_per_s"]
    return res


def kernels(epochs: int = 4, n_train: int = 61440, n_query: int = 1_000_000, seed: int = 1, batch: int = 1024):
    """Device throughput of the ensemble's kernels alone, on uniform synthetic
    inputs of the workload's shape (5 members of 40-512-512-512-1, training
    batch 1024 unless given): training steps per second and inference rows
    per second, with the BF16 FLOP counts of the GEMMs (input layer padded
    to 64)."""
    rng = np.random.default_rng(seed)
    din = sg.feature_dim("trig")
    ens = sg.MLPEnsemble(E=5, din=din, hidden=512, layers=3, seed=0)
    x = rng.uniform(-1, 1, size=(n_train, din)).astype(np.float32)
    y = rng.uniform(-1, 1, size=n_train).astype(np.float32)
    ens.fit(x[:max(4096, batch)], y[:max(4096, batch)], epochs=1, batch=batch)
    _, t = ens.fit(x, y, epochs=epochs, batch=batch)
    steps = epochs * (n_train // batch)
    flop_step = 5 * batch * 2 * (64 * 512 + 2 * 512 * 512) * 3
    q = rng.uniform(-1, 1, size=(n_query, din)).astype(np.float32)
    ens.predict(q[:65536])
    _, _, _, dev = ens.predict(q)
    flop_row = 5 * 2 * (64 * 512 + 2 * 512 * 512 + 512)
    return {"train_batch": batch, "train_steps": steps, "train_seconds": t, "train_us_per_step": t / steps * 1e6,
            "train_tflops": steps * flop_step / t / 1e12, "inference_rows": n_query,
            "inference_seconds_device": dev, "inference_tflops_device": n_query * flop_row / dev / 1e12}


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--study", action="store_true",
                    help="only the training / held-out curves of the device ensemble (raw and cos/sin inputs)")
    ap.add_argument("--encoding", default="trig", choices=sg.ENCODINGS)
    ap.add_argument("--kernels", action="store_true",
                    help="only the ensemble's training and inference throughput on synthetic inputs")
    ap.add_argument("--epochs", type=int, default=4, help="training epochs of --kernels")
    ap.add_argument("--queries", type=int, default=1_000_000, help="inference rows of --kernels")
    ap.add_argument("--batch", type=int, default=1024, help="training batch of --kernels")
    a = ap.parse_args()
    if a.kernels:
        print(json.dumps(kernels(epochs=a.epochs, n_query=a.queries, batch=a.batch)), flush=True)
    elif a.study:
        print(json.dumps(sg.learning_study(np.load(sg.LABELS_PATH)["labels"], log=log)), flush=True)
    else:
        print(json.dumps(run(encoding=a.encoding)), flush=True)
