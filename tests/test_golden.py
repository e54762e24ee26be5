"""Golden vectors: the oracle's FP32 state after a few frames of three small
scenes, committed under tests/golden/ with SHA-256 digests
(make_golden.py writes them). The CPU test holds the oracle to its
committed output; the GPU test holds the HIP path to the same bits."""
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(HERE, "golden"))
from helpers import compare  # noqa: E402

import make_golden as MG  # noqa: E402

MANIFEST = json.load(open(os.path.join(HERE, "golden", "manifest.json")))


def fixture(name):
    with np.load(os.path.join(HERE, "golden", f"{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", sorted(MG.CASES))
def test_fixture_digest(name):
    assert MG.digest(fixture(name)) == MANIFEST[name]["sha256"]


@pytest.mark.parametrize("name", sorted(MG.CASES))
def test_oracle_reproduces_golden(name):
    O = pytest.importorskip("oracle")
    got = MG.run_case(O.Oracle, name)
    want = fixture(name)
    assert sorted(got) == sorted(want)
    for k in want:
        compare(f"{name}:{k}", got[k], want[k])


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(MG.CASES))
def test_gpu_reproduces_golden(name):
    B = pytest.importorskip("blaze_amd")
    print(f"[golden] {name}", flush=True)
    got = MG.run_case(B.Solver, name)
    want = fixture(name)
    for k in want:
        compare(f"{name}:{k}", got[k], want[k])
