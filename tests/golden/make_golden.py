"""Writes the golden fixtures under tests/golden/ (run from the repo root:
``python tests/golden/make_golden.py``).

* workloads.json — the pinned inputs: every generator (z, x as hex) and every
  angle (as a float.hex string) of the 16-qubit kicked-Ising chain, the first
  two layers of the 64-qubit chain (seed 1234), one layer of the 121-qubit
  lattice, the θ_h values, the first 8 instances of the batched set
  (seed 7, metadata + first-layer angles) and the first 16 surrogate θ
  (seed 99).
* chain16_final.json — the oracle's final term set (key hex, coefficient
  float.hex) and ⟨0|O|0⟩ of the 16-qubit chain (8 layers, |c| <= 1e-5).
* prefixes.json — oracle term counts, ⟨0|O|0⟩ (float.hex) and an order-
  sensitive key/coefficient digest after the first layers of the 64-qubit
  chain (2 layers) and of the 121-qubit lattice (2 layers, θ_h index 5 and
  15), and after 4 layers of the first 3 batched instances.
* lattice121_full.json — for each of the 16 θ_h of the 121-qubit sweep, the
  live-term count after every layer and, after all 20 layers, ⟨0|O|0⟩
  (float.hex) and the digest of the final term arrays.

The outputs come from the oracle (a C restatement of the reference), so the
fixtures pin the oracle and the device path against silent drift; the
oracle itself is pinned against the reference's own tests and the
mathematics in tests/test_oracle.py.
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "paulioperators.jl_amd"), os.path.join(ROOT, "oracle")]

import numpy as np  # noqa: E402

import pyoracle as po  # noqa: E402
from pauliprop import compile_filter  # noqa: E402
from pauliprop import workloads as wl  # noqa: E402


def layer_json(L):
    return {"z": [hex(p.z) for p in L.generators], "x": [hex(p.x) for p in L.generators],
            "angles": [float(a).hex() for a in L.angles]}


def digest(z, x, c) -> str:
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(z).tobytes())
    h.update(np.ascontiguousarray(x).tobytes())
    h.update(np.ascontiguousarray(c).view(np.uint64).tobytes())
    return h.hexdigest()


def run_oracle(w, layers):
    o = po.OracleSPV(w.N, [(p.z, p.x, c) for p, c in w.observable])
    f = compile_filter(w.truncation)
    for L in w.layers[:layers]:
        o.evolve([(p.z, p.x) for p in L.generators], L.angles, strict=f)
        if w.channel:
            o.channel(*w.channel)
    return o


def workloads_json():
    k16 = wl.kicked_ising_16()
    c64 = wl.chain64(n_layers=2)
    lat = wl.lattice121(wl.theta_h_values()[5], n_layers=1)
    inst = wl.batched_instances(8, seed=7, max_depth=1)
    return {
        "kicked_ising_16": {"N": 16, "layers": [layer_json(L) for L in k16.layers]},
        "chain64_first2": {"N": 64, "layers": [layer_json(L) for L in c64.layers]},
        "lattice121_layer0_thetah5": {"N": 121, "layers": [layer_json(L) for L in lat.layers]},
        "theta_h": [float(t).hex() for t in wl.theta_h_values()],
        "batched_first8": [{"N": w.N, "meta": w.meta, "observable_z": hex(w.observable[0][0].z),
                            "layer0": layer_json(w.layers[0])} for w in inst],
        "surrogate_theta_first16": [[float(a).hex() for a in t] for t in wl.surrogate_thetas(16, seed=99)],
    }


def lattice_full():
    out = []
    for th in wl.theta_h_values():
        w = wl.lattice121(th)
        o = po.OracleSPV(w.N, [(p.z, p.x, c) for p, c in w.observable])
        f = compile_filter(w.truncation)
        sizes = []
        for L in w.layers:
            o.evolve([(p.z, p.x) for p in L.generators], L.angles, strict=f)
            o.channel(*w.channel)
            sizes.append(len(o))
        z, x, c = o.arrays()
        out.append({"theta_h": float(th).hex(), "n_by_layer": sizes, "expectation": float(o.expectation(0)).hex(),
                    "digest": digest(z, x, c)})
    return out


def main():
    with open(os.path.join(HERE, "workloads.json"), "w") as fh:
        json.dump(workloads_json(), fh)
    w = wl.kicked_ising_16()
    o = run_oracle(w, 8)
    z, x, c = o.arrays()
    with open(os.path.join(HERE, "chain16_final.json"), "w") as fh:
        json.dump({"N": 16, "n": len(c), "z": [hex(int(v)) for v in z[:, 0]], "x": [hex(int(v)) for v in x[:, 0]],
                   "c": [float(v).hex() for v in c], "expectation": float(o.expectation(0)).hex()}, fh)
    pref = {}
    cases = [("chain64_2layers", wl.chain64(n_layers=2), 2),
             ("lattice121_thetah5_2layers", wl.lattice121(wl.theta_h_values()[5], n_layers=2), 2),
             ("lattice121_thetah15_2layers", wl.lattice121(wl.theta_h_values()[15], n_layers=2), 2)]
    for w in wl.batched_instances(3, seed=7, max_depth=4):
        cases.append((f"{w.name}_4layers", w, 4))
    for name, w, nl in cases:
        o = run_oracle(w, nl)
        z, x, c = o.arrays()
        pref[name] = {"N": w.N, "layers": nl, "n": len(c), "expectation": float(o.expectation(0)).hex(),
                      "digest": digest(z, x, c)}
        print(name, len(c), file=sys.stderr)
    with open(os.path.join(HERE, "prefixes.json"), "w") as fh:
        json.dump(pref, fh, indent=1)
    with open(os.path.join(HERE, "lattice121_full.json"), "w") as fh:
        json.dump(lattice_full(), fh, indent=1)


if __name__ == "__main__":
    main()
