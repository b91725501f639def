import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")
    config.addinivalue_line("markers", "slow: long-running GPU case")


@pytest.fixture(scope="session")
def oracle():
    from oracle.oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def ref():
    from oracle.oracle import RefLib, REF_SO
    if not os.path.exists(REF_SO):
        pytest.skip("oracle/_ref not built (make -C oracle ref needs the reference checkout)")
    return RefLib()


@pytest.fixture(scope="session")
def eng():
    """The engine package with a visible GPU; GPU tests fail (not skip)
    when the library is missing or no device is visible."""
    import moonlab_amd as m
    m.lib()
    assert m.device_count() > 0, "no HIP device visible to the MI355X engine"
    return m


def load_corpus():
    import json
    idx = json.load(open(os.path.join(GOLDEN, "corpus_index.json")))["circuits"]
    z = np.load(os.path.join(GOLDEN, "corpus_ref.npz"))
    return [(k, n, z[k + ":gates"], z[k + ":amps"], z[k + ":z"]) for k, n in idx]


def random_state(n, rng):
    a = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return (a / np.linalg.norm(a)).astype(np.complex128)


ALL_GATES_1Q = [1, 2, 3, 4, 5, 6, 7, 8, 9, 14, 15, 16, 17]
ALL_GATES_2Q = [10, 11, 12, 13, 18, 19, 20, 21, 22]
ALL_GATES_3Q = [23, 24]


def random_gate_list(n, count, rng, allow3=True):
    """Random gates over the whole code table (1q, 2q, 3q)."""
    rows = []
    for _ in range(count):
        pool = ALL_GATES_1Q + (ALL_GATES_2Q if n >= 2 else []) + (ALL_GATES_3Q if allow3 and n >= 3 else [])
        t = int(rng.choice(pool))
        qs = rng.permutation(n)[:3]
        rows.append((t, int(qs[0]), int(qs[1]) if n > 1 else -1, int(qs[2]) if n > 2 else -1,
                     float(rng.uniform(-7, 7)), float(rng.uniform(-7, 7)), float(rng.uniform(-7, 7))))
    return rows
