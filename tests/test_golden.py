"""Golden fixtures (tests/golden/, written by tests/golden/make_golden.py):
the workload generators must reproduce the pinned inputs exactly, and the
oracle must reproduce the pinned outputs (term sets bit for bit, ⟨O⟩ as
float.hex, order-sensitive digests of the term arrays)."""
import json
import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden as mg  # noqa: E402


def load(name):
    with open(os.path.join(HERE, "golden", name)) as fh:
        return json.load(fh)


def test_workload_inputs_match_fixture():
    assert json.loads(json.dumps(mg.workloads_json())) == load("workloads.json")


def test_oracle_chain16_matches_fixture():
    from pauliprop import workloads as wl
    fx = load("chain16_final.json")
    o = mg.run_oracle(wl.kicked_ising_16(), 8)
    z, x, c = o.arrays()
    assert len(c) == fx["n"]
    assert [hex(int(v)) for v in z[:, 0]] == fx["z"]
    assert [hex(int(v)) for v in x[:, 0]] == fx["x"]
    assert [float(v).hex() for v in c] == fx["c"]
    assert float(o.expectation(0)).hex() == fx["expectation"]


@pytest.mark.parametrize("name", ["chain64_2layers", "lattice121_thetah5_2layers", "lattice121_thetah15_2layers",
                                  "batched_1_4layers"])
def test_oracle_prefixes_match_fixture(name):
    from pauliprop import workloads as wl
    fx = load("prefixes.json")[name]
    if name.startswith("chain64"):
        w = wl.chain64(n_layers=2)
    elif name.startswith("lattice121"):
        w = wl.lattice121(wl.theta_h_values()[5 if "thetah5" in name else 15], n_layers=2)
    else:
        w = wl.batched_instances(2, seed=7, max_depth=4)[1]
    o = mg.run_oracle(w, fx["layers"])
    z, x, c = o.arrays()
    assert len(c) == fx["n"]
    assert float(o.expectation(0)).hex() == fx["expectation"]
    assert mg.digest(z, x, c) == fx["digest"]


def test_oracle_lattice121_full_sweep_matches_fixture():
    """All 16 θ_h x 20 layers of the 121-qubit sweep: per-layer term counts,
    final ⟨O⟩ and final array digests (about 4 s of oracle time)."""
    assert mg.lattice_full() == load("lattice121_full.json")
