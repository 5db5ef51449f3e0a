import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def gpu():
    """Initialise the GPU for the session; fails (does not skip) without one:
    GPU tests are only selected with -m gpu, on a GPU machine."""
    import tagmatch_amd as tm
    tm.init(0)
    return tm


@pytest.fixture(scope="session")
def golden():
    import json
    import numpy as np

    class G:
        vectors = json.load(open(os.path.join(GOLDEN, "ref_vectors.json")))
        match = dict(np.load(os.path.join(GOLDEN, "match_small.npz")))
        partition = dict(np.load(os.path.join(GOLDEN, "partition_small.npz")))
        bitvector = dict(np.load(os.path.join(GOLDEN, "bitvector_pairs.npz")))
    return G
