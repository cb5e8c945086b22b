"""Shared test setup: repo root on sys.path, the `gpu` marker, fixture loaders."""
import gzip
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


def load_golden(name):
    with gzip.open(os.path.join(GOLDEN, name), "rt") as f:
        return json.load(f)


def hexf(s):
    return float.fromhex(s)


def golden_gates(d):
    """Gate stream of a fixture as [(kind, qubits, params)]."""
    out = []
    for kind, nq, np_, q, p in d["gates"]:
        out.append((kind, tuple(q[:nq]), tuple(hexf(x) for x in p[:np_])))
    return out


def golden_amps(d):
    import numpy as np
    a = np.array([hexf(x) for x in d["amps"]])
    return a[0::2] + 1j * a[1::2]


FUSION = {
    "fused": (1, 1, 1, 4, 0),
    "none": (0, 0, 0, 1, 0),
    "diag_only": (0, 1, 1, 4, 0),
    "win_only": (1, 0, 0, 4, 0),
    "blocked8": (1, 1, 1, 4, 8),
    "blocked10": (1, 1, 1, 4, 10),
}


@pytest.fixture(scope="session")
def qm():
    import qmetal_amd
    return qmetal_amd
