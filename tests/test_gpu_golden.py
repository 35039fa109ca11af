"""The device path against the golden fixtures (tests/golden/): the 16-qubit
chain's final term set bit for bit, and term counts, ⟨O⟩ and array digests
after the first layers of the 64-qubit chain, the 121-qubit lattice and the
first three batched instances (up to 8.2e7 live terms)."""
import json
import math
import os
import sys

import numpy as np
import pytest

from pauliprop import SparsePauliVector
from pauliprop import workloads as wl

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden as mg  # noqa: E402

pytestmark = pytest.mark.gpu


def load(name):
    with open(os.path.join(HERE, "golden", name)) as fh:
        return json.load(fh)


def run_gpu(w, layers):
    v = SparsePauliVector(w.N, [((p.z, p.x), c) for p, c in w.observable])
    for L in w.layers[:layers]:
        v.evolve_(L.generators, L.angles, truncation=w.truncation)
        if w.channel:
            getattr(v, w.channel[0] + "_channel_")(w.channel[1])
    return v


def test_chain16_matches_fixture():
    fx = load("chain16_final.json")
    v = run_gpu(wl.kicked_ising_16(), 8)
    z, x, c = v.arrays()
    assert [hex(int(q)) for q in z[:, 0]] == fx["z"]
    assert [hex(int(q)) for q in x[:, 0]] == fx["x"]
    assert [float(q).hex() for q in c] == fx["c"]
    e = float.fromhex(fx["expectation"])
    assert abs(v.expectation_value(0) - e) <= 1e-12 * max(1.0, abs(e))


@pytest.mark.parametrize("name", ["chain64_2layers", "lattice121_thetah5_2layers", "lattice121_thetah15_2layers",
                                  "batched_0_4layers", "batched_1_4layers", "batched_2_4layers"])
def test_prefixes_match_fixture(name):
    fx = load("prefixes.json")[name]
    if name.startswith("chain64"):
        w = wl.chain64(n_layers=2)
    elif name.startswith("lattice121"):
        w = wl.lattice121(wl.theta_h_values()[5 if "thetah5" in name else 15], n_layers=2)
    else:
        w = wl.batched_instances(3, seed=7, max_depth=4)[int(name.split("_")[1])]
    print(f"golden {name}: N={w.N} layers={fx['layers']}", flush=True)
    v = run_gpu(w, fx["layers"])
    assert len(v) == fx["n"]
    e = float.fromhex(fx["expectation"])
    assert abs(v.expectation_value(0) - e) <= 1e-12 * max(1.0, abs(e))
    z, x, c = v.arrays()
    assert mg.digest(z, x, c) == fx["digest"]
