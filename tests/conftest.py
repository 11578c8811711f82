import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__  # noqa: E402

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built library")


@pytest.fixture(scope="session")
def mlv():
    """The melvin_amd package (GPU tests only touch it)."""
    return __graft_entry__.load_package()


def golden(name):
    import ast

    import numpy as np

    from oracle import melvin_oracle as O

    g = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    params = ast.literal_eval(str(g["params"]))
    return g, params, O.make_params(params)


def rel(a, b):
    import numpy as np

    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))
