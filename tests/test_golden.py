"""Golden fixtures (tests/golden/, written by make_golden.py from the pinned
oracle): the oracle reproduces them on the CPU, the HIP path matches them on
the GPU."""
import json
import os

import numpy as np
import pytest

import oracle_lib as O
import polyq_amd as pq

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with open(os.path.join(HERE, name)) as f:
        return json.load(f)


def _circuit(cj):
    c = pq.Circuit(cj["n"])
    for op, a, b, ph in zip(cj["ops"], cj["q0"], cj["q1"], cj["ph"]):
        c._push(op, a, b, ph)
    return c


def test_oracle_reproduces_amplitude_fixtures():
    for name in ("amplitudes_config1_seed1.json", "amplitudes_config2_seed1.json"):
        g = _load(name)
        c = _circuit(g["circuit"])
        s = O.OSim(O.OCircuit.from_circuit(c), "cached")
        s.set_threads(4)
        for x, y, re, im in zip(g["x"], g["y"], g["re"], g["im"]):
            assert abs(s.amplitude(x, y) - complex(re, im)) < 1e-12


def test_oracle_reproduces_branch_fixtures():
    for case in _load("branches_small.json")["cases"]:
        c = _circuit(case["circuit"])
        s = O.OSim(O.OCircuit.from_circuit(c), "cached")
        assert list(s.stats()) == case["stats"]
        zero, r, k8, kt = s.branch_dump(case["x"], case["y"], np.asarray(case["s"], np.uint64))
        assert zero.tolist() == case["zero"] and r.tolist() == case["r"] and k8.tolist() == case["k8"] and kt.tolist() == case["ktot"]


def test_product_host_stats_match_fixtures():
    for case in _load("branches_small.json")["cases"]:
        assert list(pq.Simulator(_circuit(case["circuit"]), strategy="cached").stats()) == case["stats"]
    g = _load("amplitudes_config2_seed1.json")
    assert list(pq.Simulator(_circuit(g["circuit"]), strategy="cached").stats()) == g["stats"]


@pytest.mark.gpu
def test_gpu_matches_amplitude_fixtures():
    for name in ("amplitudes_config1_seed1.json", "amplitudes_config2_seed1.json"):
        g = _load(name)
        c = _circuit(g["circuit"])
        want = np.array(g["re"]) + 1j * np.array(g["im"])
        scale = np.max(np.abs(want))
        for form, rel in (("plain_fp64", 1e-12), ("z8", 1e-11), ("branch_walk", 1e-6)):
            got = pq.Simulator(c, strategy="cached", form=form).amplitudes(g["x"], g["y"])
            assert np.max(np.abs(got - want)) <= rel * scale, (name, form)


@pytest.mark.gpu
def test_gpu_matches_branch_fixtures():
    for case in _load("branches_small.json")["cases"]:
        s = pq.Simulator(_circuit(case["circuit"]), strategy="cached")
        zero, k8, kt, r = s.branch_dump(case["x"], case["y"], np.asarray(case["s"], np.uint64))
        assert zero.tolist() == case["zero"]
        live = [i for i, z in enumerate(case["zero"]) if not z]
        assert [int(k8[i]) for i in live] == [case["k8"][i] for i in live]
        assert [int(kt[i]) for i in live] == [case["ktot"][i] for i in live]
        assert all(r == case["r"][i] for i in live)
