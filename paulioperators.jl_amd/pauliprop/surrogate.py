"""Learned surrogate of ⟨O⟩(θ) with an ensemble-disagreement gate.

The exact method is the propagation engine itself: ⟨0|O(θ)|0⟩ of the
36-qubit chain (10 layers, per-layer shared RZZ and RX angles, θ ∈ [-π, π]^20,
observable Z_18, |c| <= 1e-6) computed by SparsePauliBatch on the GPU.  The
surrogate is an ensemble of E MLPs (20 -> 512 -> 512 -> 512 -> 1, GELU)
trained with Adam on labels the engine produced; at query time a θ whose
ensemble standard deviation exceeds ``tol`` is sent back to exact
propagation.  Training and inference run in libpauliprop's BF16
matrix-core kernels (pp_mlp_*).  No part of this has a counterpart in the
reference.
"""
from __future__ import annotations

import ctypes
import math
import os
import time
from typing import Callable, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import ArgumentError, check
from . import workloads as wl

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
LABELS_PATH = os.path.join(DATA_DIR, "chain36_labels.npz")
N_TRAIN, N_HELDOUT = 65536, 4096

_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)


ENCODINGS = ("raw", "trig")


def features(thetas, encoding: str = "raw") -> np.ndarray:
    """Network input, float32: "raw" θ / π ([n][20]); "trig" cos θ and sin θ
    of every angle ([n][40]: ⟨O⟩ is 2π-periodic in each angle)."""
    t = np.asarray(thetas, dtype=np.float64)
    if encoding == "raw":
        return (t / math.pi).astype(np.float32)
    if encoding == "trig":
        return np.concatenate([np.cos(t), np.sin(t)], axis=1).astype(np.float32)
    raise ArgumentError(f"encoding must be one of {ENCODINGS}")


def feature_dim(encoding: str) -> int:
    return 20 if encoding == "raw" else 40


class MLPEnsemble:
    """E independent MLPs din -> hidden^layers -> 1 on the device."""

    def __init__(self, E: int = 5, din: int = 20, hidden: int = 512, layers: int = 3, seed: int = 0):
        self.E, self.din, self.hidden, self.layers = E, din, hidden, layers
        h = ctypes.c_void_p()
        check(_lib.lib().pp_mlp_create(E, din, hidden, layers, ctypes.byref(h)))
        self._h = h
        c = ctypes.c_int64()
        check(_lib.lib().pp_mlp_param_count(self._h, ctypes.byref(c)))
        self.count = c.value
        self.set_params(self.init_params(seed))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().pp_mlp_destroy(h)
            except Exception:
                pass
            self._h = None

    def init_params(self, seed: int) -> np.ndarray:
        """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights and biases per layer,
        one independent draw per member."""
        rng = np.random.default_rng(seed)
        out = np.empty((self.E, self.count), dtype=np.float32)
        H = self.hidden
        for e in range(self.E):
            parts = []
            for l in range(self.layers):
                kin = self.din if l == 0 else H
                bound = 1.0 / math.sqrt(kin)
                parts.append(rng.uniform(-bound, bound, size=H * kin))
                parts.append(rng.uniform(-bound, bound, size=H))
            bound = 1.0 / math.sqrt(H)
            parts.append(rng.uniform(-bound, bound, size=H))
            parts.append(rng.uniform(-bound, bound, size=1))
            out[e] = np.concatenate(parts)
        return out

    def set_params(self, params: np.ndarray):
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.shape != (self.E, self.count):
            raise ArgumentError(f"params must have shape {(self.E, self.count)}")
        check(_lib.lib().pp_mlp_set_params(self._h, p.ctypes.data_as(_fp)))

    def get_params(self) -> np.ndarray:
        p = np.empty((self.E, self.count), dtype=np.float32)
        check(_lib.lib().pp_mlp_get_params(self._h, p.ctypes.data_as(_fp)))
        return p

    def gradients(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.empty((self.E, self.count), dtype=np.float32)
        loss = np.empty(self.E, dtype=np.float32)
        check(_lib.lib().pp_mlp_gradients(self._h, x.ctypes.data_as(_fp), t.ctypes.data_as(_fp), len(t),
                                          g.ctypes.data_as(_fp), loss.ctypes.data_as(_fp)))
        return g, loss

    def fit(self, x, y, epochs: int = 60, batch: int = 1024, lr: float = 1e-3, lr_min: float = 1e-5,
            seed: int = 1, log: Optional[Callable[[str], None]] = None,
            callback: Optional[Callable[[int, np.ndarray], None]] = None, every: int = 10):
        """Adam with a cosine learning-rate decay from lr to lr_min; each member
        sees its own shuffle of the data every epoch.  callback(epoch, losses)
        runs after every `every`-th epoch and the last (outside the timed
        training).  Returns the per-epoch mean losses [epochs][E] and the wall
        time of the training epochs."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        n = len(y)
        check(_lib.lib().pp_mlp_load_data(self._h, x.ctypes.data_as(_fp), y.ctypes.data_as(_fp), n))
        rng = np.random.default_rng(seed)
        hist = []
        t_train = 0.0
        for ep in range(epochs):
            t0 = time.perf_counter()
            perm = np.stack([rng.permutation(n) for _ in range(self.E)]).astype(np.int32)
            rate = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * ep / max(epochs - 1, 1)))
            loss = np.empty(self.E, dtype=np.float32)
            check(_lib.lib().pp_mlp_train_epoch(self._h, perm.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                                batch, ctypes.c_float(rate), loss.ctypes.data_as(_fp)))
            t_train += time.perf_counter() - t0
            hist.append(loss.copy())
            if log and (ep % 10 == 0 or ep == epochs - 1):
                log(f"epoch {ep}: lr {rate:.2e} loss {np.array2string(loss, precision=3)}")
            if callback and (ep % every == 0 or ep == epochs - 1):
                callback(ep, loss.copy())
        return np.array(hist), t_train

    def predict(self, x, members: bool = False):
        """(mean, std, members or None, device seconds) for x[n][din]."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        n = len(x)
        mean = np.empty(n, dtype=np.float64)
        sd = np.empty(n, dtype=np.float64)
        mem = np.empty((self.E, n), dtype=np.float32) if members else None
        secs = ctypes.c_double()
        check(_lib.lib().pp_mlp_predict(self._h, x.ctypes.data_as(_fp), n,
                                        mem.ctypes.data_as(_fp) if members else None,
                                        mean.ctypes.data_as(_dp), sd.ctypes.data_as(_dp), ctypes.byref(secs)))
        return mean, sd, mem, secs.value


# --- exact method: the propagation engine ----------------------------------------

def _chain36_layer_arrays(thetas, li: int, W: int):
    """Packed steps of layer li for every θ (the circuits share generators;
    RZZ angle θ[2 li] on the 35 bonds, RX angle θ[2 li + 1] on the 36 sites)."""
    from .spv import _pack
    N = 36
    gens = [(p.z, p.x) for p in wl.surrogate_chain36_layers([0.0] * 20).layers[li].generators]
    B = len(thetas)
    S = len(gens)
    gz = np.zeros((S, B, W), dtype=np.uint64)
    gx = np.zeros((S, B, W), dtype=np.uint64)
    for s, (z, x) in enumerate(gens):
        gz[s, :] = _pack(z, W)
        gx[s, :] = _pack(x, W)
    nb = N - 1
    cz = np.array([math.cos(t[2 * li]) for t in thetas])
    sz = np.array([math.sin(t[2 * li]) for t in thetas])
    cx = np.array([math.cos(t[2 * li + 1]) for t in thetas])
    sx = np.array([math.sin(t[2 * li + 1]) for t in thetas])
    cv = np.empty((S, B))
    sv = np.empty((S, B))
    cv[:nb], sv[:nb] = cz, sz
    cv[nb:], sv[nb:] = cx, sx
    return gz, gx, cv, sv, np.ones((S, B), dtype=np.uint8)


def chain36_batch_expectations(thetas: Sequence[Sequence[float]], batch: int = 256,
                               log: Optional[Callable[[str], None]] = None) -> np.ndarray:
    """⟨0|O(θ)|0⟩ for every θ by batched GPU propagation (SparsePauliBatch):
    `batch` circuits per batch, one evolve! per layer per circuit."""
    from .batch import SparsePauliBatch
    from .truncation import CoeffTruncation
    out = np.empty(len(thetas), dtype=np.float64)
    obs = wl.surrogate_chain36_layers([0.0] * 20).observable
    for b0 in range(0, len(thetas), batch):
        th = [list(t) for t in thetas[b0:b0 + batch]]
        bt = SparsePauliBatch([36] * len(th), [[((p.z, p.x), c) for p, c in obs]] * len(th))
        for li in range(10):
            bt.evolve_packed(_chain36_layer_arrays(th, li, bt.W), truncation=CoeffTruncation(1e-6))
        out[b0:b0 + len(th)] = bt.expectation_values()
        if log:
            log(f"labels {b0 + len(th)}/{len(thetas)}")
    return out


def chain36_plain_expectation(theta: Sequence[float]) -> float:
    """One θ at a time: a SparsePauliVector and one evolve! per layer."""
    from .spv import SparsePauliVector
    w = wl.surrogate_chain36_layers(list(theta))
    v = SparsePauliVector(w.N, [((p.z, p.x), c) for p, c in w.observable])
    for L in w.layers:
        v.evolve_(L.generators, L.angles, truncation=w.truncation)
    return v.expectation_value(0)


class GatedSurrogate:
    """Ensemble mean where the ensemble agrees, exact propagation elsewhere.

    A θ is served by the network when its ensemble standard deviation is at
    most ``threshold`` (a disagreement, in units of ⟨O⟩'s spread over the
    members; it is not the error tolerance).  ``calibrate`` sets it from
    labelled θ the network was not trained on: the largest s such that every
    calibration point with std <= s is within ``tol`` of its exact value
    (ties at s count as above it), accepted only when at least
    ``min_points`` calibration points lie at or below s (no miss among m
    points bounds the miss rate on new θ below 3/m at 95% confidence);
    otherwise nothing is served and every query goes to exact propagation.
    (Ensemble members that all collapse to the same wrong answer agree
    closely, so disagreement alone is not a safe gate.)"""

    def __init__(self, ensemble: MLPEnsemble, exact: Callable = chain36_batch_expectations, tol: float = 1e-3,
                 encoding: str = "raw", min_points: int = 256):
        self.ens, self.exact, self.tol, self.encoding = ensemble, exact, tol, encoding
        self.min_points = min_points
        self.threshold = -1.0  # uncalibrated: serve nothing

    def calibrate(self, thetas, labels):
        mean, sd, _, _ = self.ens.predict(features(thetas, self.encoding))
        err = np.abs(mean - np.asarray(labels))
        order = np.argsort(sd, kind="stable")
        bad = np.nonzero(err[order] > self.tol)[0]
        k = bad[0] if len(bad) else len(order)  # points [0, k) of the std order are all within tol
        self.threshold = -1.0
        while k > 0:
            s = float(sd[order[k - 1]])
            if k == len(order) or sd[order[k]] > s:
                break
            k -= 1  # a tie across the first miss: step below it
        if k >= max(1, self.min_points):
            self.threshold = float(sd[order[k - 1]])
        return {"threshold": self.threshold, "calibration_points": int(len(order)), "within_tol_prefix": int(k),
                "min_points": int(self.min_points), "calibration_rmse": float(np.sqrt(np.mean(err ** 2)))}

    def serves(self, sd) -> np.ndarray:
        return sd <= self.threshold if self.threshold >= 0 else np.zeros(len(sd), dtype=bool)

    def evaluate(self, thetas):
        thetas = np.asarray(thetas, dtype=np.float64)
        t0 = time.perf_counter()
        mean, sd, _, dev_s = self.ens.predict(features(thetas, self.encoding))
        t1 = time.perf_counter()
        fall = ~self.serves(sd)
        vals = mean.copy()
        if fall.any():
            vals[fall] = self.exact([list(t) for t in thetas[fall]])
        t2 = time.perf_counter()
        return vals, fall, {"surrogate_s": t1 - t0, "surrogate_device_s": dev_s, "exact_s": t2 - t1}


# --- learnability of the labels ----------------------------------------------------

def learning_study(labels, encodings=ENCODINGS, capacity_n: int = 4096, capacity_epochs: int = 300,
                   sizes=(4096, 16384, 61440), epochs: int = 60, n_eval: int = 4096, E: int = 5,
                   log: Optional[Callable[[str], None]] = None) -> dict:
    """Training and held-out curves of the device ensemble on the chain36
    labels.  capacity: the first `capacity_n` training θ for
    `capacity_epochs` epochs (can the network fit them at all?); sizes: the
    first n training θ for `epochs` epochs (does the fit carry over to θ it
    has not seen?).  Held-out RMSE on the first `n_eval` held-out θ; the
    reference line is the labels' standard deviation (predicting the mean)."""
    tr, ho = training_thetas()
    ytr, yho = labels[:len(tr)], labels[len(tr):len(tr) + n_eval]
    ho = ho[:n_eval]
    out = {"label_var_train": float(np.var(ytr)), "label_std_heldout": float(np.std(yho)), "runs": []}

    def one(enc, n, ep_count, every):
        ens = MLPEnsemble(E=E, din=feature_dim(enc), hidden=512, layers=3, seed=0)
        xf, hf = features(tr[:n], enc), features(ho, enc)
        curve = []

        def cb(ep, loss):
            mean, _, _, _ = ens.predict(hf)
            curve.append({"epoch": ep, "train_loss": float(np.mean(loss)),
                          "heldout_rmse": float(np.sqrt(np.mean((mean - yho) ** 2)))})

        _, secs = ens.fit(xf, ytr[:n], epochs=ep_count, batch=1024, lr=1e-3, lr_min=1e-5, callback=cb, every=every)
        mean, _, _, _ = ens.predict(xf)
        run = {"encoding": enc, "n_train": n, "epochs": ep_count, "train_seconds": secs,
               "train_rmse_final": float(np.sqrt(np.mean((mean - ytr[:n]) ** 2))), "curve": curve}
        if log:
            log(f"{enc} n={n}: train loss {curve[-1]['train_loss']:.5f}, held-out rmse {curve[-1]['heldout_rmse']:.4f}")
        out["runs"].append(run)

    for enc in encodings:
        one(enc, capacity_n, capacity_epochs, max(1, capacity_epochs // 10))
        for n in sizes:
            one(enc, n, epochs, max(1, epochs // 6))
    return out


# --- labels --------------------------------------------------------------------------

def training_thetas():
    """The 65,536 training and 4,096 held-out θ: the first 69,632 draws of the
    seed-99 stream, in that order."""
    th = np.array(wl.surrogate_thetas(N_TRAIN + N_HELDOUT, seed=99), dtype=np.float64)
    return th[:N_TRAIN], th[N_TRAIN:]


def query_thetas(count: int = 1_000_000, seed: int = 100) -> np.ndarray:
    """Query θ for inference throughput: uniform [-π, π]^20 (numpy, seed 100)."""
    return np.random.default_rng(seed).uniform(-math.pi, math.pi, size=(count, 20))


def generate_labels(path: str = LABELS_PATH, start: int = 0, stop: Optional[int] = None, batch: int = 512,
                    budget_s: float = float("inf"), log: Optional[Callable[[str], None]] = None):
    """Label θ rows [start, stop) of training + held-out θ by GPU propagation,
    resuming from the .npz at `path` (key `labels`, NaN where not computed)
    and saving it after every batch; stops after `budget_s` seconds."""
    tr, ho = training_thetas()
    allth = np.concatenate([tr, ho])
    stop = len(allth) if stop is None else stop
    labels = np.full(len(allth), np.nan)
    if os.path.exists(path):
        labels = np.load(path)["labels"]
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    t0 = time.perf_counter()
    todo = [i for i in range(start, stop) if np.isnan(labels[i])]
    for b0 in range(0, len(todo), batch):
        idx = todo[b0:b0 + batch]
        t1 = time.perf_counter()
        labels[idx] = chain36_batch_expectations(allth[idx], batch=batch)
        np.savez_compressed(path, labels=labels)
        if log:
            log(f"labels: {b0 + len(idx)}/{len(todo)} done, batch {time.perf_counter() - t1:.2f} s")
        if time.perf_counter() - t0 > budget_s:
            break
    return labels


if __name__ == "__main__":
    import argparse
    import sys
    ap = argparse.ArgumentParser(description="generate chain36 surrogate labels by GPU propagation")
    ap.add_argument("--start", type=int, default=0)
    ap.add_argument("--stop", type=int, default=None)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--out", default=LABELS_PATH)
    ap.add_argument("--budget", type=float, default=float("inf"), help="seconds")
    a = ap.parse_args()
    generate_labels(a.out, a.start, a.stop, a.batch, a.budget, log=lambda m: print(m, file=sys.stderr, flush=True))
