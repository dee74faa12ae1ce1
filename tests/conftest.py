"""Shared fixtures: the drop-in library, the C restatement and (where built)
the reference build, all loaded side by side with RTLD_LOCAL."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "planck-levels_amd"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    return O.load_oracle()


@pytest.fixture(scope="session")
def ref():
    import oracle as O
    r = O.load_ref()
    if r is None:
        pytest.skip("reference build oracle/_ref not present")
    return r


@pytest.fixture(scope="session")
def product():
    """The drop-in library driven through the generic libsharp wrapper."""
    import oracle as O
    import sharp_hip  # noqa: F401  (raises if the library is missing)
    return O.SharpLike(sharp_hip.LIB_PATH, "sharp_")


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.linalg.norm(b.ravel())
    return np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0)
